"""Writes tests/golden/zest_golden.npz from the upstream zest library.

Run in a container that holds the zest sources, after
`make -f oracle/Makefile ref` has compiled them into oracle/_ref. The
fixtures are inputs and zest's outputs for the GLQ transforms
(zt::GLQTransformer, include/zernike_glq_transformer.hpp:640-679, and
st::GLQTransformer, include/sh_glq_transformer.hpp:632-668); inputs come from
the portable generator zest_amd.synth, so they can be regenerated anywhere.

Large grids are stored as sampled values at seeded positions plus
checksums (sum and sum of squares), to keep the file small.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from zest_amd import synth  # noqa: E402

CONV_ZT = [(zn, sn, sp) for zn in (0, 1) for sn in (0, 1) for sp in (-1, 1)]
CONV_ST = [(sn, sp) for sn in (0, 1) for sp in (-1, 1)]


def sample_idx(size, count, seed):
    return (synth.splitmix64(seed, count) % np.uint64(size)).astype(np.int64)


def main():
    if oracle.REF is None:
        raise SystemExit("oracle/_ref/libzest_ref.so is not built")
    out = {}
    # small Zernike cases, every convention: full grids and coefficients
    for order in (1, 2, 5, 6, 9):
        for (zn, sn, sp) in (CONV_ZT if order < 9 else [(1, 0, -1), (0, 1, 1)]):
            key = f"zt_{order}_{zn}{sn}{sp + 1}"
            z = oracle.ZestZernike(order, zn, sn, sp)
            c = synth.zernike_coeffs(order, 100 + order)
            g_rand = synth.normal(200 + order, int(np.prod(oracle.zt_shape(order)))).reshape(oracle.zt_shape(order))
            out[key + "_coeffs"] = c
            out[key + "_bwd"] = z.backward(c)
            out[key + "_grid"] = g_rand
            out[key + "_fwd"] = z.forward(g_rand)
    # truncation: grid order 9 with expansion orders 5 and 12
    z = oracle.ZestZernike(9, 1, 0, -1)
    g = synth.normal(300, int(np.prod(oracle.zt_shape(9)))).reshape(oracle.zt_shape(9))
    sentinel = np.full((oracle.zt_num_coeffs(12), 2), 7.25)
    out["zt_trunc_grid"] = g
    out["zt_trunc_fwd5"] = z.forward(g, exp_order=5)
    out["zt_trunc_fwd12"] = z.forward(g, exp_order=12, out=sentinel.copy())
    c12 = synth.zernike_coeffs(12, 301)
    out["zt_trunc_c12"] = c12
    out["zt_trunc_bwd12"] = z.backward(c12, exp_order=12)
    # order 32, seed 0 (the reference's CPU-runnable case): grid = zest backward
    c = synth.zernike_coeffs(32, 0)
    z = oracle.ZestZernike(32, 1, 0, -1)
    g = z.backward(c)
    idx = sample_idx(g.size, 2000, 32)
    out["zt32_coeffs"] = c
    out["zt32_grid_idx"] = idx
    out["zt32_grid_val"] = g.ravel()[idx]
    out["zt32_grid_sums"] = np.array([g.sum(), (g * g).sum()])
    out["zt32_fwd"] = z.forward(g)
    # order 128, seeds 0..3: sampled backward values, forward of the grid
    z = oracle.ZestZernike(128, 1, 0, -1)
    for seed in range(4):
        c = synth.zernike_coeffs(128, seed)
        g = z.backward(c)
        idx = sample_idx(g.size, 1000, 128 + seed)
        out[f"zt128_{seed}_grid_idx"] = idx
        out[f"zt128_{seed}_grid_val"] = g.ravel()[idx]
        out[f"zt128_{seed}_grid_sums"] = np.array([g.sum(), (g * g).sum()])
        f = z.forward(g)
        cidx = sample_idx(f.shape[0], 1000, 256 + seed)
        out[f"zt128_{seed}_fwd_idx"] = cidx
        out[f"zt128_{seed}_fwd_val"] = f[cidx]
    # spherical harmonics, both layouts, every convention
    for order in (1, 2, 5, 7, 12):
        for layout in (0, 1):
            for (sn, sp) in (CONV_ST if order < 12 else [(0, -1)]):
                key = f"st_{order}_{layout}_{sn}{sp + 1}"
                s = oracle.ZestSH(order, sn, sp, layout)
                c = synth.sh_coeffs(order, 400 + order)
                shape = oracle.st_shape(order, layout)
                g_rand = synth.normal(500 + order, int(np.prod(shape))).reshape(shape)
                out[key + "_coeffs"] = c
                out[key + "_bwd"] = s.backward(c)
                out[key + "_grid"] = g_rand
                out[key + "_fwd"] = s.forward(g_rand)
    # order 512, seed 1: sampled values
    s = oracle.ZestSH(512, 0, -1, 0)
    c = synth.sh_coeffs(512, 1)
    g = s.backward(c)
    idx = sample_idx(g.size, 2000, 512)
    out["st512_grid_idx"] = idx
    out["st512_grid_val"] = g.ravel()[idx]
    out["st512_grid_sums"] = np.array([g.sum(), (g * g).sum()])
    f = s.forward(g)
    cidx = sample_idx(f.shape[0], 2000, 513)
    out["st512_fwd_idx"] = cidx
    out["st512_fwd_val"] = f[cidx]
    # Gauss-Legendre nodes and weights (table path < 70 nodes, Bogaert >= 70)
    for n in (1, 2, 7, 32, 33, 69, 70, 71, 129, 136, 512):
        zz, ww = oracle.ref_gl_nodes_weights(n)
        out[f"gl_{n}_nodes"] = zz
        out[f"gl_{n}_weights"] = ww
    path = os.path.join(HERE, "zest_golden.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}: {len(out)} arrays, {os.path.getsize(path) / 1e3:.0f} kB")


if __name__ == "__main__":
    main()
