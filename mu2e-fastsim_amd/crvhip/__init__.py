"""crvhip: the CRV response chain of crvfastsim on MI355X (gfx950) through HIP kernels behind a C-ABI."""
