"""Parameter block of the learned subgrid closure (wf_solver_set_closure).

The closure is an MLP 12 -> 64 -> 64 -> 1 (GELU hidden layers, softplus
output) that gives nu_t per node; the solver evaluates it with BF16 MFMA on
the GPU (wildfires_amd/csrc/hip/wf_closure.h).  This module packs, unpacks,
saves and loads the WF_CLOSURE_NPARAMS float32 block whose layout
include/wildfires.h documents:

    lo[12], hi[12], mean[12], inv_std[12], W1[64][12], b1[64], W2[64][64],
    b2[64], w3[64], b3, out_scale

Inputs, in order: ux uy uz vx vy vz wx wy wz, |grad T|, Delta, Y.
"""
import numpy as np

NIN = 12
NH = 64
FEATURES = ("ux", "uy", "uz", "vx", "vy", "vz", "wx", "wy", "wz", "grad_T", "delta", "Y")
_LAYOUT = (("lo", (NIN,)), ("hi", (NIN,)), ("mean", (NIN,)), ("inv_std", (NIN,)),
           ("W1", (NH, NIN)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)),
           ("w3", (NH,)), ("b3", ()), ("out_scale", ()))
NPARAMS = sum(int(np.prod(shape)) for _, shape in _LAYOUT)


def pack(**parts):
    """float32 parameter block from its named parts (every name of the layout)."""
    out = np.zeros(NPARAMS, dtype=np.float32)
    o = 0
    for name, shape in _LAYOUT:
        n = int(np.prod(shape))
        v = np.asarray(parts[name], dtype=np.float32).reshape(-1)
        if v.size != n:
            raise ValueError("closure part %s: %d values, %d expected" % (name, v.size, n))
        out[o:o + n] = v
        o += n
    return out


def unpack(block):
    """Named parts (float32 arrays; b3 and out_scale as floats) of a block."""
    block = np.asarray(block, dtype=np.float32).reshape(-1)
    if block.size != NPARAMS:
        raise ValueError("closure block: %d values, %d expected" % (block.size, NPARAMS))
    parts, o = {}, 0
    for name, shape in _LAYOUT:
        n = int(np.prod(shape))
        v = block[o:o + n].reshape(shape)
        parts[name] = float(v) if shape == () else v.copy()
        o += n
    return parts


def save(path, block, **meta):
    """The block (and optional metadata arrays) as an .npz file."""
    np.savez(path, closure=np.asarray(block, dtype=np.float32), **meta)


def load(path):
    """The block of a file written by save (no pickled objects are read)."""
    with np.load(path, allow_pickle=False) as f:
        block = np.asarray(f["closure"], dtype=np.float32)
    unpack(block)  # checks the size
    return block


def random_block(seed=0, lo=None, hi=None, mean=None, inv_std=None, out_scale=1e-3, w_scale=0.5):
    """A seeded untrained network (Gaussian weights) with the given input range
    and normalisation (defaults: the whole real line, identity)."""
    rng = np.random.default_rng(seed)
    big = np.float32(3.0e38)
    return pack(lo=np.full(NIN, -big) if lo is None else lo, hi=np.full(NIN, big) if hi is None else hi,
                mean=np.zeros(NIN) if mean is None else mean, inv_std=np.ones(NIN) if inv_std is None else inv_std,
                W1=rng.normal(0.0, w_scale, (NH, NIN)), b1=rng.normal(0.0, 0.1, NH),
                W2=rng.normal(0.0, w_scale / np.sqrt(NH / NIN), (NH, NH)), b2=rng.normal(0.0, 0.1, NH),
                w3=rng.normal(0.0, w_scale / np.sqrt(NH / NIN), NH), b3=0.0, out_scale=out_scale)
