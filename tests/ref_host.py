"""Runs the reference's own host code (oracle/_ref/libref_host.so, compiled from io.cu, mathfunc.cu,
preprocess.cu, bb_core.cu, bb_utils.cu and variables.cu) on one run directory and dumps its global
state as .npy files.  Test infrastructure: it is started as a child process by the tests, because
the reference keeps its state in process globals and reads its files relative to the working
directory (main.cu:49-88).

    python tests/ref_host.py CASE_DIR OUT_DIR [--bb] [--write]

Steps, in the order of main.cu:56-88:
  read_inputs("input")  (io.cu:12)    -> scalars.npy
  read_BB_inputs(...)   (io.cu:337)   -> bb_*.npy                  (with --bb)
  read_observers(...)   (io.cu:106)   -> rObs.npy
  read_surf(nSurf)      (io.cu:206)   -> raw_{X,N,PR}.npy, E.npy
  preprocess_imper()    (preprocess.cu:228) -> pre_{X,N,V,PR}.npy
  writeTimeHistory / writeBBoutput (io.cu:137, 702) on seeded arrays  (with --write)
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_PATH = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "libref_host.so")


def seeded_outputs(nTime, nObs, nf2):
    """Arrays handed to the writers by both the reference and the product writer."""
    g = np.random.default_rng(1234)
    p = [(g.standard_normal(nTime * nObs) * 10.0 ** g.integers(-6, 3, nTime * nObs)).astype(np.float32)
         for _ in range(3)]
    oaspl = (40.0 + 30.0 * g.random(nObs)).astype(np.float32)
    spl = (-20.0 + 100.0 * g.random(nObs * nf2 * 7)).astype(np.float32)
    return p, oaspl, spl


def main(argv):
    case, out = os.path.abspath(argv[1]), os.path.abspath(argv[2])
    bb = "--bb" in argv
    write = "--write" in argv
    os.makedirs(out, exist_ok=True)
    lib = C.CDLL(REF_PATH)
    g = lambda ty, name: ty.in_dll(lib, name)
    fptr = lambda name: C.POINTER(C.c_float).in_dll(lib, name)
    iptr = lambda name: C.POINTER(C.c_int).in_dll(lib, name)
    arr = lambda p, n, dt: np.ctypeslib.as_array(p, shape=(n,)).astype(dt).copy()

    os.chdir(case)
    lib._Z11read_inputsPKc(b"input")
    if bb:
        lib._Z14read_BB_inputsPKc(b"input_BB")
    lib._Z14read_observersPKc(b"obs.dat")
    nSurf = g(C.c_int, "nSurf").value
    nTime = g(C.c_int, "nTime").value
    nObs = g(C.c_int, "nObs").value
    lib._Z9read_surfi(C.c_int(nSurf))
    E = arr(iptr("E"), nSurf, np.int64)
    q = int(E.sum()) * nTime
    sc = {n: g(C.c_float, n).value for n in ("a0", "rhoRef", "pRef", "oM", "dPsi", "Minf", "Mref",
                                              "dTau", "CFDscale", "OBSscale")}
    sc.update({n: g(C.c_int, n).value for n in ("nSurf", "nTime", "nObs", "periodic", "lowpass",
                                                 "impermeable", "BB_noise")})
    vinf = np.ctypeslib.as_array((C.c_float * 3).in_dll(lib, "vInf")).copy()
    np.save(os.path.join(out, "scalars.npy"), np.array([sc[k] for k in sorted(sc)], np.float64))
    np.save(os.path.join(out, "scalar_names.npy"), np.array(sorted(sc)))
    np.save(os.path.join(out, "vInf.npy"), vinf)
    np.save(os.path.join(out, "E.npy"), E)
    np.save(os.path.join(out, "rObs.npy"), arr(fptr("rObs"), 3 * nObs, np.float32))
    for nm, w in (("X", 4), ("N", 6), ("PR", 2)):
        np.save(os.path.join(out, f"raw_{nm}.npy"), arr(fptr(nm + "ham"), w * q, np.float32))
    lib._Z16preprocess_imperv()
    for nm, w in (("X", 4), ("N", 6), ("V", 6), ("PR", 2)):
        np.save(os.path.join(out, f"pre_{nm}.npy"), arr(fptr(nm + "ham"), w * q, np.float32))

    if bb:
        nS = g(C.c_int, "nSurf_BB").value
        nsect = arr(iptr("nsect"), nS, np.int64)
        tot = int(nsect.sum())
        np.save(os.path.join(out, "bb_nsect.npy"), nsect)
        np.save(os.path.join(out, "bb_xyzcR.npy"), arr(fptr("xyzcR"), 4 * tot, np.float32))
        np.save(os.path.join(out, "bb_bbdata.npy"), arr(fptr("bbdata"), 7 * tot * nTime, np.float32))
        for nm in ("span", "omega", "psi_offset"):
            np.save(os.path.join(out, f"bb_{nm}.npy"), arr(fptr(nm), nS, np.float32))
        for nm in ("rotate", "translate"):
            np.save(os.path.join(out, f"bb_{nm}.npy"), arr(fptr(nm), 3 * nS, np.float32))
        for nm in ("ccw", "trip", "nsect_t", "nmach_t", "naoa_t", "nre_t"):
            np.save(os.path.join(out, f"bb_{nm}.npy"), arr(iptr(nm), nS, np.int64))
        nst = np.load(os.path.join(out, "bb_nsect_t.npy"))
        nm_ = np.load(os.path.join(out, "bb_nmach_t.npy"))
        na = np.load(os.path.join(out, "bb_naoa_t.npy"))
        nr = np.load(os.path.join(out, "bb_nre_t.npy"))
        np.save(os.path.join(out, "bb_sect_t.npy"), arr(fptr("sect_t"), int(nst.sum()), np.float32))
        np.save(os.path.join(out, "bb_mach_t.npy"), arr(fptr("mach_t"), int(nm_.sum()), np.float32))
        np.save(os.path.join(out, "bb_aoa_t.npy"), arr(fptr("aoa_t"), int(na.sum()), np.float32))
        np.save(os.path.join(out, "bb_re_t.npy"), arr(fptr("re_t"), int(nr.sum()), np.float32))
        np.save(os.path.join(out, "bb_bbdata_t.npy"),
                arr(fptr("bbdata_t"), int((nst * nm_ * na * nr * 7).sum()), np.float32))
        nf2 = g(C.c_int, "nf2").value
        np.save(os.path.join(out, "bb_ff.npy"), arr(fptr("ff"), nf2, np.float32))
        np.save(os.path.join(out, "bb_ints.npy"), np.array([nS, g(C.c_int, "t_del").value, nf2], np.int64))

    if write:
        nf2 = g(C.c_int, "nf2").value if bb else 1
        (pT, pL, pA), oaspl, spl = seeded_outputs(nTime, nObs, nf2)
        os.chdir(out)
        fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
        lib._Z16writeTimeHistoryPfS_S_(fp(pT), fp(pL), fp(pA))
        if bb:
            # writeBBoutput frees both of its arguments (io.cu:795): hand it malloc'd copies
            libc = C.CDLL(None)
            libc.malloc.restype = C.c_void_p
            libc.malloc.argtypes = [C.c_size_t]

            def cbuf(a):
                p = libc.malloc(a.nbytes)
                C.memmove(p, a.ctypes.data, a.nbytes)
                return C.cast(p, C.POINTER(C.c_float))
            lib._Z13writeBBoutputPfS_(cbuf(oaspl), cbuf(spl))
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
