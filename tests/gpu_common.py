"""Shared helpers of the GPU parity tests: the parity bar and the oracle runs.

Bar (floating point, scalar_t = float on the GPU vs the oracle's double build):
displacement, rotation and strains within 1e-4 px / rad, gamma within 1e-5
relative on subsets where both converge, identical status on >= 99.9 % of
subsets."""
import os

import numpy as np

import dice_amd as dh
import oracle_py as o

U_TOL = 1e-4
GAMMA_RTOL = 1e-5
STATUS_AGREE = 0.999
SHAPE_FIELDS = ("DISP_X", "DISP_Y", "ROT_Z", "NS_XX", "NS_YY", "SS_XY")


def oracle_threads():
    return max(1, min(16, os.cpu_count() or 1))


def to_oracle(p):
    q = o.Params()
    for name, _ in dh.Params._fields_:
        setattr(q, name, getattr(p, name))
    return q


def run_oracle(p, ref, de, cx, cy, subset_size, fields=None, conformal=None, first_frame=True,
               precision="double"):
    f = o.new_fields(cx, cy) if fields is None else np.ascontiguousarray(fields).copy()
    o.correlate(to_oracle(p), ref, de, f, subset_size=subset_size, conformal=conformal, first_frame=first_frame,
                precision=precision, threads=oracle_threads())
    return f


def compare(fg, fo, u_tol=U_TOL, gamma_rtol=GAMMA_RTOL, status_agree=STATUS_AGREE, label=""):
    """Status agreement, then shape parameters and gamma on the subsets both runs converged."""
    sg, so = fg[dh.F["STATUS"]], fo[o.F["STATUS"]]
    agree = float(np.mean(sg == so))
    if agree < status_agree:
        pairs, counts = np.unique(np.stack([sg[sg != so], so[sg != so]]), axis=1, return_counts=True)
        detail = ", ".join("gpu %d / oracle %d: %d" % (a, b, c) for (a, b), c in zip(pairs.T, counts))
        raise AssertionError("%s status agreement %.5f (%s)" % (label, agree, detail))
    both = np.isin(sg, dh.CONVERGED_STATUSES) & np.isin(so, dh.CONVERGED_STATUSES)
    for name in SHAPE_FIELDS:
        d = np.abs(fg[dh.F[name]] - fo[o.F[name]])[both]
        assert d.size == 0 or d.max() <= u_tol, "%s %s max diff %.3g" % (label, name, d.max())
    g, go = fg[dh.F["GAMMA"]][both], fo[o.F["GAMMA"]][both]
    rel = np.abs(g - go) / np.maximum(np.abs(go), 1e-30)
    assert rel.size == 0 or rel.max() <= gamma_rtol, "%s gamma max rel diff %.3g" % (label, rel.max())
    return agree, both


def csr_take(conf, idx):
    """The CSR pixel lists of the subsets idx (in that order)."""
    off, px, py, act = conf
    cnt = off[idx + 1] - off[idx]
    new_off = np.zeros(len(idx) + 1, np.int32)
    np.cumsum(cnt, out=new_off[1:])
    sel = np.concatenate([np.arange(off[i], off[i + 1]) for i in idx])
    return new_off, px[sel], py[sel], act[sel]
