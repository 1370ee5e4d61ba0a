"""Configuration of the CRV response chain: the reference's dataclasses plus the closed-form optics.

Field names, defaults and meaning follow the reference (crvfastsim): SiPMConfig (response/sipm.py:26-47),
WaveformConfig (reco/waveform.py:29-44), RecoConfig (reco/pulsefit.py:26-46), SectorCuts / FinderConfig
(reco/coinc.py:29-61), ResponseConfig (response/photons.py:45-53, yield scale and per-channel variation only) and
ChainConfig (chain.py:61-84). Additions of this build: OpticsConfig (39 PE/MeV at 1 m, two-exponential attenuation,
fibre speed, emission delay) replacing the lookup tables, the 128-sample waveform window, the photon map as data,
the Philox seed, and the device precision ("fp64" plain stages or "fp32" batched chain).
"""
from dataclasses import dataclass, field

import numpy as np

PERIOD = 12.5
CHAN_MULT = 32768


@dataclass(frozen=True)
class SiPMConfig:
    n_pixels_x: int = 40
    n_pixels_y: int = 40
    overvoltage: float = 3.0
    time_constant: float = 13.3
    capacitance: float = 8.84e-14
    p1: float = 0.607
    p2: float = 2.7
    cross_talk: float = 0.05
    thermal_rate: float = 0.0
    trap0_prob: float = 0.0
    trap1_prob: float = 0.0
    trap0_lifetime: float = 5.0
    trap1_lifetime: float = 50.0
    inactive: tuple = tuple((x, y) for x in range(18, 22) for y in range(18, 22))
    pixel_id: str = "lrint"


@dataclass(frozen=True)
class WaveformConfig:
    single_pe_file: str = "synthetic"
    single_pe_precision: float = 0.5
    single_pe_stretch: float = 1.047
    single_pe_max_time: float = 100.0
    single_pe_reference_charge: float = 2.652e-13
    min_voltage: float = 0.0275
    noise: float = 4.0e-4
    samples_zs: int = 12
    adc_conversion: float = 2300.0
    pedestal: int = 100
    min_adc: int = 0
    max_adc: int = 4095


@dataclass(frozen=True)
class RecoConfig:
    min_adc_difference: float = 40.0
    default_beta: float = 19.0
    min_beta: float = 5.0
    max_beta: float = 40.0
    max_time_difference: float = 15.0
    min_pulse_height_ratio: float = 0.7
    max_pulse_height_ratio: float = 1.5
    le_time_factor: float = 0.985
    samples_before: int = 4
    samples_after: int = 3
    pedestal_undershoot_threshold: float = 6.0
    use_time_offset_db: bool = True
    time_offset_scale: float = 1.0
    ignore_channels: bool = True
    max_pulses: int = 8
    lm_iterations: int = 40
    limit_tolerance: float = 0.01
    knife_edge_adc: float = 1.0


@dataclass(frozen=True)
class SectorCuts:
    pe_threshold: float
    max_dt_adjacent: float
    max_dt: float
    min_slope: float
    max_slope: float
    max_slope_diff: float
    coincidence_layers: int
    min_cluster_pes: float
    initial_max_distance: float
    final_max_distance: float


# the synthetic detector's one sector: 8 PE, 10 ns adjacent window, 3 of 4 layers within 10 ns, slopes +-6, max slope
# difference 3, initial / final max distance 250 / 0 mm (the reference's EXTRACTED_V04 T1 cuts with 3 layers)
SYNTH_CUTS = SectorCuts(8.0, 10.0, 10.0, -6.0, 6.0, 3.0, 3, 0.0, 250.0, 0.0)


@dataclass(frozen=True)
class FinderConfig:
    fiber_signal_speed: float = 140.0
    max_time_interval: float = 20.0
    time_offset: float = 34.5
    compensate_status_bits: tuple = (0, 1, 2)
    ignore_secondary_pulses: bool = True
    big_cluster_threshold: int = 50000
    cuts: dict = field(default_factory=lambda: {"CRV": SYNTH_CUTS})
    add_opposite_side_pulses: bool = True


@dataclass(frozen=True)
class SectorConfig:
    """Lookup-table optics of one sector (reference response/photons.py:31-35): table name, reflector, photons/MeV."""
    table: str
    reflector: int
    yield_per_mev: float


# the stock extracted-detector sectors (reference response/photons.py:38-43, Offline epilog_extracted_v05.fcl)
EXTRACTED_V05 = {"EX": SectorConfig("6000_0", 0, 39794.0), "T1": SectorConfig("2370_1", -1, 28610.0),
                 "T2": SectorConfig("3200_0", 0, 28610.0),
                 **{f"M{i}": SectorConfig("2370_1", 1, 28610.0) for i in range(1, 9)}}


@dataclass(frozen=True)
class ResponseConfig:
    """photon_yield_scale multiplies every arrival probability (the reference's default is 0.94; the closed-form optics
    of this build are calibrated at 1.0). sectors, speed_ratio and atten_ratio configure the lookup-table optics
    (charges(..., tabs=...)); the closed-form optics ignore them."""
    photon_yield_scale: float = 1.0
    photon_yield_variation_scale: float = 1.0
    sectors: dict = field(default_factory=lambda: dict(EXTRACTED_V05))
    speed_ratio: dict = field(default_factory=dict)
    atten_ratio: dict = field(default_factory=dict)


@dataclass(frozen=True)
class OpticsConfig:
    """Closed-form light yield: mu = edep * Y * A(d) / A(d_ref) per SiPM, A(d) = w_s e^{-d/l_s} + w_l e^{-d/l_l}."""
    pe_per_mev: float = 39.0
    ref_distance: float = 1000.0      # mm
    w_short: float = 0.2
    l_short: float = 300.0            # mm
    w_long: float = 0.8
    l_long: float = 4500.0            # mm
    half_length: float = 2550.0       # mm, readout ends at z = -H (side 0) and +H (side 1)
    inv_speed: float = 1.0 / 140.0    # ns/mm along the fibre
    tau: float = 7.4                  # ns, emission delay time constant


@dataclass(frozen=True)
class ChainConfig:
    response: ResponseConfig = field(default_factory=ResponseConfig)
    optics: OpticsConfig = field(default_factory=OpticsConfig)
    finder: FinderConfig = field(default_factory=FinderConfig)
    sipm: SiPMConfig = field(default_factory=SiPMConfig)
    window: tuple = (0.0, 1e5)
    noise: bool = True
    waveform: WaveformConfig = field(default_factory=WaveformConfig)
    pulse_reco: RecoConfig = field(default_factory=RecoConfig)
    feb_phase_width: float = PERIOD
    feb_phase_key: str = "roc_feb"
    window_samples: int = 128
    photon_map: np.ndarray = field(default_factory=lambda: np.ones((40, 40)))
    background: object = None         # crvhip.background.BackgroundModel: out-of-time pulses added before clustering
    background_scale: float = 1.0     # multiplies the background model's net rate
    backend: str = "hip"
    precision: str = "fp64"

    def __hash__(self):
        return id(self)
