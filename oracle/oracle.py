"""ORACLE — test infrastructure only.

ctypes front-end of the FP64 CPU restatement in ``oracle/oracle.cpp``.  Only the
tests, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` may
import this module; the product package never does.
"""
import ctypes as C
import json
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}

D = C.POINTER(C.c_double)
I = C.POINTER(C.c_int)
U8 = C.POINTER(C.c_ubyte)


def build():
    subprocess.run(["make", "-s", "-C", _HERE, "-j2"], check=True)


def _lib(fast=False):
    name = "liboracle_fast.so" if fast else "liboracle.so"
    if name not in _LIBS:
        path = os.path.join(_HERE, name)
        if not os.path.exists(path):
            build()
        lib = C.CDLL(path)
        lib.orc_create.restype = C.c_void_p
        lib.orc_create.argtypes = [C.c_char_p]
        lib.orc_destroy.argtypes = [C.c_void_p]
        lib.orc_last_error.restype = C.c_char_p
        lib.orc_bjerrum_length.restype = C.c_double
        lib.orc_bjerrum_length.argtypes = [C.c_double, C.c_double]
        lib.orc_run_translate.argtypes = [C.c_void_p, C.c_long, C.c_uint, C.c_int, C.c_double, I, I, D, D, D, U8, D, D]
        lib.orc_replay_translate.argtypes = [C.c_void_p, C.c_long, I, I, D, D, D, U8, D, D]
        lib.orc_run_philox_translate.argtypes = [C.c_void_p, C.c_long, C.c_ulonglong, C.c_uint, C.c_ulonglong,
                                                 C.c_double, I, D, D, U8]
        lib.orc_mc_stats.argtypes = [C.c_void_p, C.POINTER(C.c_long), C.POINTER(C.c_long), D]
        lib.orc_run_philox_moves.argtypes = [C.c_void_p, C.c_long, C.c_ulonglong, C.c_uint, C.c_ulonglong, C.c_double,
                                             C.c_int, I, I, C.POINTER(C.c_uint), D, D, I, I, C.POINTER(C.c_long),
                                             C.c_long, I, D, D, D, U8]
        lib.orc_molecule_excluded.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, I]
        lib.orc_group_mass_center.argtypes = [C.c_void_p, C.c_int, C.c_int, D]
        _LIBS[name] = lib
    return _LIBS[name]


def _p(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t)) if a is not None else None


def _check(lib, rc):
    if rc != 0:
        raise RuntimeError(lib.orc_last_error().decode())


class System:
    """One system (both Metropolis states) of the restated reference path."""

    def __init__(self, config, fast=False):
        self.lib = _lib(fast)
        text = config if isinstance(config, str) else json.dumps(config)
        h = self.lib.orc_create(text.encode())
        if not h:
            raise RuntimeError(self.lib.orc_last_error().decode())
        self.h = C.c_void_p(h)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_destroy(self.h)
            self.h = None

    def set_particles(self, xyz, charge, atom_id):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        charge = np.ascontiguousarray(charge, dtype=np.float64)
        atom_id = np.ascontiguousarray(atom_id, dtype=np.int32)
        self.n = len(charge)
        _check(self.lib, self.lib.orc_set_particles(self.h, self.n, _p(xyz), _p(charge), _p(atom_id, C.c_int)))

    def set_groups(self, begin, size, molid, mass_center=None):
        begin = np.ascontiguousarray(begin, dtype=np.int32)
        size = np.ascontiguousarray(size, dtype=np.int32)
        molid = np.ascontiguousarray(molid, dtype=np.int32)
        cm = None if mass_center is None else np.ascontiguousarray(mass_center, dtype=np.float64)
        self.max_group = int(size.max()) if len(size) else 0
        _check(self.lib, self.lib.orc_set_groups(self.h, len(begin), _p(begin, C.c_int), _p(size, C.c_int),
                                                 _p(molid, C.c_int), _p(cm)))

    def init(self):
        e = C.c_double()
        _check(self.lib, self.lib.orc_init(self.h, C.byref(e)))
        return e.value

    @property
    def num_terms(self):
        return self.lib.orc_num_terms(self.h)

    def term_names(self):
        out = []
        for i in range(self.num_terms):
            buf = C.create_string_buffer(64)
            _check(self.lib, self.lib.orc_term_name(self.h, i, buf, 64))
            out.append(buf.value.decode())
        return out

    def total_energy(self, trial=False):
        t = np.zeros(self.num_terms)
        tot = C.c_double()
        _check(self.lib, self.lib.orc_total_energy(self.h, int(trial), _p(t), C.byref(tot)))
        return tot.value, t

    def energy_change(self, change, moved_index, new_xyz):
        """Per-term energies (trial, accepted) of a Change; see fmc_energy_change."""
        gi, al, it, off, idx = change_arrays(change)
        moved_index = np.ascontiguousarray(moved_index, dtype=np.int32)
        new_xyz = np.ascontiguousarray(new_xyz, dtype=np.float64)
        en = np.zeros(self.num_terms)
        eo = np.zeros(self.num_terms)
        _check(self.lib, self.lib.orc_energy_change(
            self.h, int(change.get("everything", False)), len(gi), _p(gi, C.c_int), _p(al, C.c_int),
            _p(it, C.c_int), _p(off, C.c_int), _p(idx, C.c_int), len(moved_index), _p(moved_index, C.c_int),
            _p(new_xyz), _p(en), _p(eo)))
        return en, eo

    def run_translate(self, nmoves, seed, molid=0, default_dp=0.0):
        nt = self.num_terms
        rec = dict(group=np.zeros(nmoves, np.int32), index=np.zeros(nmoves, np.int32),
                   xyz=np.zeros((nmoves, 3)), u=np.zeros(nmoves), du=np.zeros(nmoves),
                   accepted=np.zeros(nmoves, np.uint8), e_new=np.zeros((nmoves, nt)), e_old=np.zeros((nmoves, nt)))
        _check(self.lib, self.lib.orc_run_translate(
            self.h, nmoves, seed, molid, default_dp, _p(rec["group"], C.c_int), _p(rec["index"], C.c_int),
            _p(rec["xyz"]), _p(rec["u"]), _p(rec["du"]), _p(rec["accepted"], C.c_ubyte), _p(rec["e_new"]),
            _p(rec["e_old"])))
        return rec

    def run_philox_translate(self, nmoves, seed, replica=0, move0=0, beta=1.0):
        """Metropolis run with the device-resident Philox proposal scheme."""
        rec = dict(index=np.zeros(nmoves, np.int32), xyz=np.zeros((nmoves, 3)), du=np.zeros(nmoves),
                   accepted=np.zeros(nmoves, np.uint8))
        _check(self.lib, self.lib.orc_run_philox_translate(
            self.h, nmoves, seed, replica, move0, beta, _p(rec["index"], C.c_int), _p(rec["xyz"]), _p(rec["du"]),
            _p(rec["accepted"], C.c_ubyte)))
        return rec

    def run_philox_moves(self, nmoves, seed, kinds, replica=0, move0=0, beta=1.0, cap=None):
        """Metropolis run over a mix of move kinds (0 atom translate, 1 molecule
        translate+rotate, 2 pivot) with the device-resident Philox proposal scheme.
        ``kinds``: list of dicts with keys kind, molid, weight, dp, dprot.  Records per
        move: kind, group, moved atoms (offsets into index/xyz), trial mass centre,
        du and decision."""
        nk = len(kinds)
        kind = np.array([k["kind"] for k in kinds], np.int32)
        molid = np.array([k["molid"] for k in kinds], np.int32)
        weight = np.array([k.get("weight", 1) for k in kinds], np.uint32)
        dp = np.array([k.get("dp", 0.0) for k in kinds])
        dprot = np.array([k.get("dprot", 0.0) for k in kinds])
        cap = cap if cap is not None else nmoves * max(1, self.max_group)
        rec = dict(kind=np.zeros(nmoves, np.int32), group=np.zeros(nmoves, np.int32),
                   offsets=np.zeros(nmoves + 1, np.int64), index=np.zeros(cap, np.int32), xyz=np.zeros((cap, 3)),
                   cm=np.zeros((nmoves, 3)), du=np.zeros(nmoves), accepted=np.zeros(nmoves, np.uint8))
        _check(self.lib, self.lib.orc_run_philox_moves(
            self.h, nmoves, seed, replica, move0, beta, nk, _p(kind, C.c_int), _p(molid, C.c_int),
            _p(weight, C.c_uint), _p(dp), _p(dprot), _p(rec["kind"], C.c_int), _p(rec["group"], C.c_int),
            _p(rec["offsets"], C.c_long), cap, _p(rec["index"], C.c_int), _p(rec["xyz"]), _p(rec["cm"]),
            _p(rec["du"]), _p(rec["accepted"], C.c_ubyte)))
        used = int(rec["offsets"][-1])
        rec["index"] = rec["index"][:used]
        rec["xyz"] = rec["xyz"][:used]
        return rec

    def excluded(self, molid, i, j):
        e = C.c_int()
        _check(self.lib, self.lib.orc_molecule_excluded(self.h, molid, i, j, C.byref(e)))
        return bool(e.value)

    def mass_center(self, group, trial=False):
        cm = np.zeros(3)
        _check(self.lib, self.lib.orc_group_mass_center(self.h, int(trial), group, _p(cm)))
        return cm

    def replay_translate(self, group, index, xyz, u):
        n = len(u)
        nt = self.num_terms
        out = dict(du=np.zeros(n), accepted=np.zeros(n, np.uint8), e_new=np.zeros((n, nt)), e_old=np.zeros((n, nt)))
        group = np.ascontiguousarray(group, np.int32)
        index = np.ascontiguousarray(index, np.int32)
        xyz = np.ascontiguousarray(xyz, np.float64)
        u = np.ascontiguousarray(u, np.float64)
        _check(self.lib, self.lib.orc_replay_translate(
            self.h, n, _p(group, C.c_int), _p(index, C.c_int), _p(xyz), _p(u), _p(out["du"]),
            _p(out["accepted"], C.c_ubyte), _p(out["e_new"]), _p(out["e_old"])))
        return out

    def positions(self, trial=False):
        xyz = np.zeros((self.n, 3))
        _check(self.lib, self.lib.orc_get_positions(self.h, int(trial), _p(xyz)))
        return xyz

    def stats(self):
        m, a, s = C.c_long(), C.c_long(), C.c_double()
        _check(self.lib, self.lib.orc_mc_stats(self.h, C.byref(m), C.byref(a), C.byref(s)))
        return m.value, a.value, s.value

    def pair_energy_exact(self, ta, qa, tb, qb, r2):
        out = C.c_double()
        _check(self.lib, self.lib.orc_pair_energy_exact(self.h, ta, C.c_double(qa), tb, C.c_double(qb),
                                                        C.c_double(r2), C.byref(out)))
        return out.value

    def splined_table(self, ta, tb, cap=4096):
        """nonbonded_splined table of a type pair: knots, coefficients, rmin2, rmax2, hardsphere."""
        knots = np.zeros(cap)
        coeff = np.zeros(6 * cap)
        nk, hs = C.c_int(), C.c_int()
        r0, r1 = C.c_double(), C.c_double()
        _check(self.lib, self.lib.orc_splined_table(self.h, ta, tb, cap, _p(knots), _p(coeff), C.byref(nk),
                                                    C.byref(r0), C.byref(r1), C.byref(hs)))
        n = nk.value
        return dict(knots=knots[:n], coeff=coeff[:6 * (n - 1)], rmin2=r0.value, rmax2=r1.value,
                    hardsphere=bool(hs.value))

    def pair_energy(self, ta, qa, tb, qb, r2):
        out = C.c_double()
        _check(self.lib, self.lib.orc_pair_energy(self.h, ta, C.c_double(qa), tb, C.c_double(qb),
                                                  C.c_double(r2), C.byref(out)))
        return out.value


def change_arrays(change):
    groups = change.get("groups", [])
    gi = np.array([g["group_index"] for g in groups], np.int32)
    al = np.array([int(g.get("all", False)) for g in groups], np.int32)
    it = np.array([int(g.get("internal", False)) for g in groups], np.int32)
    off = np.zeros(len(groups) + 1, np.int32)
    idx = []
    for k, g in enumerate(groups):
        rel = list(g.get("relative_atom_indices", []))
        idx += rel
        off[k + 1] = off[k] + len(rel)
    return gi, al, it, off, np.array(idx, np.int32)


def bjerrum_length(temperature, epsr):
    return _lib().orc_bjerrum_length(temperature, epsr)


def ewald_parts(coulomb, temperature, box, xyz, q):
    lib = _lib()
    xyz = np.ascontiguousarray(xyz, np.float64)
    q = np.ascontiguousarray(q, np.float64)
    nk = C.c_int()
    s, su, r = C.c_double(), C.c_double(), C.c_double()
    _check(lib, lib.orc_ewald_parts(json.dumps(coulomb).encode(), C.c_double(temperature), C.c_double(box),
                                    len(q), _p(xyz), _p(q), C.byref(nk), C.byref(s), C.byref(su), C.byref(r)))
    return nk.value, s.value, su.value, r.value


def ewald_kvectors(coulomb, temperature, box, cap=200000):
    lib = _lib()
    out = np.zeros((cap, 4))
    nk = C.c_int()
    _check(lib, lib.orc_ewald_kvectors(json.dumps(coulomb).encode(), C.c_double(temperature), C.c_double(box),
                                       cap, _p(out), C.byref(nk)))
    return out[: nk.value]


def spline_test(which, coulomb=None, utol=1e-5, ftol=-1.0, xmin=0.0, xmax=1.0, xs=()):
    lib = _lib()
    cap = 2000
    knots = np.zeros(cap)
    coeff = np.zeros(cap * 6)
    nk = C.c_int()
    xs = np.ascontiguousarray(xs, np.float64)
    ys = np.zeros(len(xs))
    cj = json.dumps(coulomb or {}).encode()
    _check(lib, lib.orc_spline_test(which, cj, C.c_double(utol), C.c_double(ftol), C.c_double(xmin),
                                    C.c_double(xmax), cap, _p(knots), _p(coeff), C.byref(nk), len(xs), _p(xs),
                                    _p(ys)))
    n = nk.value
    return knots[:n], coeff[: 6 * (n - 1)], ys


def philox(ctr, key):
    lib = _lib()
    c = (C.c_uint * 4)(*ctr)
    k = (C.c_uint * 2)(*key)
    o = (C.c_uint * 4)()
    lib.orc_philox(c, k, o)
    return list(o)
