"""Writes ref_mollib_vectors.npz: input/output vectors of the reference's own compiled
Src/mollib.F90 routines (Random, CardinalBSpline, ErfLocal), from oracle/_ref/libref_mollib.so
(built by `make -C oracle ref` from the reference's source where it lies).

Run from the repository root:  python tests/golden/make_ref_mollib_vectors.py
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
SEEDS = (-1, -4711, -12345, 0, 7)
NDRAW = 2000
ORDERS = tuple(range(2, 13))


def ref_lib():
    path = os.path.join(ROOT, "oracle", "_ref", "libref_mollib.so")
    L = C.CDLL(path)
    dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
    L.ref_random.argtypes = [C.c_int, C.POINTER(C.c_int32), ip, C.POINTER(C.c_double), dp]
    L.ref_bspline.argtypes = [C.c_int, C.c_int, dp, ip, dp, dp]
    L.ref_erf.argtypes = [C.c_int, dp, dp]
    return L


def ref_random(L, seed, n, state=(-1, -1), am=0.0):
    idum = C.c_int32(seed)
    st = np.array(state, dtype=np.int32)
    a = C.c_double(am)
    out = np.zeros(n)
    L.ref_random(n, C.byref(idum), st, C.byref(a), out)
    return out, idum.value, st, a.value


def ref_bspline(L, order, pos):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    k = np.zeros(pos.size, dtype=np.int32)
    v = np.zeros((pos.size, order))
    d = np.zeros((pos.size, order))
    L.ref_bspline(order, pos.size, pos, k, v, d)
    return k, v, d


def main():
    L = ref_lib()
    out = {}
    for s in SEEDS:
        v, idum, st, am = ref_random(L, s, NDRAW)
        out[f"random_{s}"] = v
        out[f"random_{s}_end"] = np.array([idum, st[0], st[1]], dtype=np.int64)
        out[f"random_{s}_am"] = np.array([am])
    rng = np.random.default_rng(2024)
    pos = np.concatenate([rng.uniform(-40.0, 140.0, 40), np.arange(-2.0, 2.01, 0.25)])
    out["bspline_pos"] = pos
    for n in ORDERS:
        k, v, d = ref_bspline(L, n, pos)
        out[f"bspline_{n}_kmax"], out[f"bspline_{n}_val"], out[f"bspline_{n}_der"] = k, v, d
    x = np.linspace(0.0, 6.0, 241)
    e = np.zeros_like(x)
    L.ref_erf(x.size, x, e)
    out["erf_x"], out["erf_val"] = x, e
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_mollib_vectors.npz"), **out)


if __name__ == "__main__":
    main()
