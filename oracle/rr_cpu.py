"""ctypes wrapper of oracle/_build/librr_cpu.so (the C/OpenMP restatement of rr_full).

TEST INFRASTRUCTURE ONLY: used by tests/ as a cross-check of the numpy oracle and
by bench.py's CPU-baseline leg.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, '_build', 'librr_cpu.so')
_lib = None


def build():
    subprocess.run(['make', '-s', '-C', HERE], check=True)
    return LIB


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = ctypes.CDLL(LIB)
        _lib.rrcpu_evaluate.restype = ctypes.c_int
        _lib.rrcpu_weights_3d.restype = None
    return _lib


def _p(x):
    return x.ctypes.data_as(ctypes.c_void_p)


def weights_3d(nk, kmin, kmax, ze, ng):
    lib = load()
    ze = np.ascontiguousarray(ze, float)
    W = np.zeros((nk, ng, ze.size))
    lib.rrcpu_weights_3d(ctypes.c_int(nk), ctypes.c_double(kmin), ctypes.c_double(kmax), _p(ze), ctypes.c_int(ze.size),
                         ctypes.c_int(ng), _p(W))
    return W


LAW_IDS = {'uniform': 0, 'linear': 1, 'quadratic': 2, 'quadratic-tri': 3, 'nonlinear': 4, 'general': 5,
           'square_root': 6, 'logarithmic': 7, 'exponential': 8, 'power-2': 9, 'power-2-pm': 10}


class CpuRR:
    """RoadRunnerModel.evaluate restated in C/OpenMP, FP64."""

    def __init__(self, law='quadratic', klims=(0.005, 0.5), nk=256, nzin=20, nzlimb=20, zcut=0.7, ng=100):
        from . import rr_oracle as o
        self.law = LAW_IDS[law]
        self.klims, self.nk, self.ng = klims, nk, ng
        self.ze, zm = o.create_z_grid(zcut, nzin, nzlimb)
        self.mu = o.mu_nodes(zm)
        self.dk = (klims[1] - klims[0]) / nk
        self.dg = float(np.linspace(0.0, 1.0 - 1e-7, ng)[1])
        self.W = weights_3d(nk, klims[0], klims[1], self.ze, ng)

    def set_data(self, time, lcids=None, pbids=None, nsamples=1, exptimes=0.0, epids=None):
        self.time = np.ascontiguousarray(time, float)
        npt = self.time.size
        self.lcids = np.ascontiguousarray(lcids if lcids is not None else np.zeros(npt), np.int32)
        self.nlc = int(np.unique(self.lcids).size)
        self.pbids = np.ascontiguousarray(pbids if pbids is not None else np.zeros(self.nlc), np.int32)
        self.npb = int(np.unique(self.pbids).size)
        self.epids = np.ascontiguousarray(epids if epids is not None else np.zeros(self.nlc), np.int32)
        self.nsamples = np.zeros(self.nlc, np.int32)
        self.nsamples[:] = nsamples
        self.exptimes = np.zeros(self.nlc)
        self.exptimes[:] = exptimes

    def evaluate(self, k, ldc, t0, p, a, i, e, w, nthreads=0):
        lib = load()
        k = np.ascontiguousarray(np.atleast_2d(k), float)
        t0 = np.ascontiguousarray(np.atleast_2d(t0), float)
        npv = k.shape[0]
        arr = [np.ascontiguousarray(np.broadcast_to(np.asarray(x, float), (npv,))) for x in (p, a, i, e, w)]
        ldc = np.ascontiguousarray(np.broadcast_to(np.asarray(ldc, float).reshape(npv, -1, np.shape(ldc)[-1]),
                                                   (npv, self.npb, np.shape(ldc)[-1])))
        flux = np.zeros((npv, self.time.size))
        ci = ctypes.c_int
        lib.rrcpu_evaluate(ci(self.time.size), _p(self.time), ci(self.nlc), _p(self.lcids), ci(self.npb),
                           _p(self.pbids), _p(self.epids), _p(self.nsamples), _p(self.exptimes), ci(npv),
                           ci(k.shape[1]), _p(k), ci(t0.shape[1]), _p(t0), *[_p(x) for x in arr], ci(self.law),
                           ci(ldc.shape[2]), _p(ldc), ci(self.mu.size), _p(self.mu), ci(self.nk), ci(self.ng),
                           ctypes.c_double(self.klims[0]), ctypes.c_double(self.klims[1]), ctypes.c_double(self.dk),
                           ctypes.c_double(self.dg), _p(self.W), _p(self.ze), _p(flux), ci(nthreads))
        return flux
