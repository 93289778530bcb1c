"""faunus_amd — MI355X energy engine for Faunus's Metropolis Monte Carlo.

Python front-end over the C ABI declared in ``include/faunus_amd.h`` (the library
``faunus_amd/libfaunus_amd.so`` built from ``faunus_amd/csrc``).  There is no CPU
fallback: every energy is computed by the HIP kernels, and loading fails loudly
when the library is missing.
"""
import ctypes as C
import json
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfaunus_amd.so")
_lib = None

D = C.POINTER(C.c_double)
I = C.POINTER(C.c_int)
U8 = C.POINTER(C.c_ubyte)

STATE_ACCEPTED = 0
STATE_TRIAL = 1

# every entry point of include/faunus_amd.h with its ctypes signature
SIGNATURES = {
    "fmc_last_error": (C.c_char_p, []),
    "fmc_device_count": (C.c_int, [I]),
    "fmc_create": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "fmc_destroy": (C.c_int, [C.c_void_p]),
    "fmc_set_particles": (C.c_int, [C.c_void_p, C.c_int, D, D, I]),
    "fmc_set_groups": (C.c_int, [C.c_void_p, C.c_int, I, I, I, D]),
    "fmc_init": (C.c_int, [C.c_void_p, D]),
    "fmc_num_terms": (C.c_int, [C.c_void_p, I]),
    "fmc_term_name": (C.c_int, [C.c_void_p, C.c_int, C.c_char_p, C.c_int]),
    "fmc_stage_mass_center": (C.c_int, [C.c_void_p, C.c_int, D]),
    "fmc_get_mass_center": (C.c_int, [C.c_void_p, C.c_int, D]),
    "fmc_stage_trial": (C.c_int, [C.c_void_p, C.c_int, I, D]),
    "fmc_delta_u": (C.c_int, [C.c_void_p, C.c_void_p, D, D]),
    "fmc_energy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, D, D]),
    "fmc_sync": (C.c_int, [C.c_void_p, C.c_int]),
    "fmc_total_energy": (C.c_int, [C.c_void_p, D, D]),
    "fmc_check_q": (C.c_int, [C.c_void_p, D, D]),
    "fmc_get_positions": (C.c_int, [C.c_void_p, D]),
    "fmc_replay_translate": (C.c_int, [C.c_void_p, C.c_long, I, I, D, D, D, U8, D, D]),
    "fmc_replay_moves": (C.c_int, [C.c_void_p, C.c_long, I, C.POINTER(C.c_long), I, D, D, U8, D, D, U8]),
    "fmc_run_translate": (C.c_int, [C.c_void_p, C.c_long, I, I, D, D, D, U8]),
    "fmc_set_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "fmc_get_timing": (C.c_int, [C.c_void_p, D, C.POINTER(C.c_long)]),
    "fmc_batch_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "fmc_batch_destroy": (C.c_int, [C.c_void_p]),
    "fmc_batch_add_replica": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int, D, D, I, C.c_int, I, I, I, C.c_double,
                                        C.c_ulonglong, C.c_int, I]),
    "fmc_batch_add_replica_moves": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int, D, D, I, C.c_int, I, I, I, C.c_double,
                                              C.c_ulonglong, C.c_int, C.c_int, I, I, C.POINTER(C.c_uint), D, D, I]),
    "fmc_batch_run": (C.c_int, [C.c_void_p, C.POINTER(C.c_longlong), C.c_int]),
    "fmc_batch_run_split": (C.c_int, [C.c_void_p, C.c_longlong, C.c_int, C.c_int, C.c_int]),
    "fmc_batch_run_split_static": (C.c_int, [C.c_void_p, C.c_longlong, C.c_int, C.c_int]),
    "fmc_batch_capture_interval": (C.c_int, [C.c_void_p, C.POINTER(C.c_longlong), C.c_ulonglong, C.c_ulonglong]),
    "fmc_batch_replay": (C.c_int, [C.c_void_p, C.c_int]),
    "fmc_batch_run_ladders": (C.c_int, [C.c_void_p, C.POINTER(C.c_longlong), I, C.c_ulonglong, C.c_ulonglong]),
    "fmc_batch_swap_counts": (C.c_int, [C.c_void_p, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
    "fmc_batch_autotune": (C.c_int, [C.c_void_p, C.c_longlong, C.c_int, I, I, D, I]),
    "fmc_batch_run_pipe": (C.c_int, [C.c_void_p, C.c_longlong, C.c_int, C.c_int, C.c_int]),
    "fmc_batch_split_error": (C.c_int, [C.c_void_p, I]),
    "fmc_batch_sync": (C.c_int, [C.c_void_p, D]),
    "fmc_batch_set_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "fmc_batch_swap": (C.c_int, [C.c_void_p, C.c_ulonglong, C.c_ulonglong, C.POINTER(C.c_longlong),
                                 C.POINTER(C.c_longlong)]),
    "fmc_batch_replica_info": (C.c_int, [C.c_void_p, C.c_int, D, D, C.POINTER(C.c_ulonglong),
                                         C.POINTER(C.c_ulonglong), I]),
    "fmc_batch_positions": (C.c_int, [C.c_void_p, C.c_int, D]),
    "fmc_batch_trace": (C.c_int, [C.c_void_p, C.c_int, C.c_int, I, U8, D, D]),
    "fmc_batch_total_energy": (C.c_int, [C.c_void_p, C.c_int, D, D]),
    "fmc_batch_sample_energy": (C.c_int, [C.c_void_p, C.c_longlong]),
    "fmc_batch_set_cell_divisions": (C.c_int, [C.c_void_p, C.c_int]),
    "fmc_batch_set_cell_capacity": (C.c_int, [C.c_void_p, C.c_int]),
    "fmc_batch_get_replica_state": (C.c_int, [C.c_void_p, C.c_int, D, D,
                                              C.POINTER(C.c_ulonglong), C.POINTER(C.c_ulonglong)]),
    "fmc_batch_set_replica_state": (C.c_int, [C.c_void_p, C.c_int, D, D, C.c_ulonglong,
                                              C.c_ulonglong]),
    "fmc_batch_energy_stats": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_ulonglong), D, D]),
    "fmc_batch_rdf": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int,
                                C.POINTER(C.c_ulonglong)]),
    "fmc_batch_check_q": (C.c_int, [C.c_void_p, C.c_int, D, D]),
    "fmc_copy_probe": (C.c_int, [C.c_longlong, C.c_int, D]),
    "fmc_coulomb_spline": (C.c_int, [C.c_void_p, C.c_int, D, D, I]),
    "fmc_num_kvectors": (C.c_int, [C.c_void_p, I]),
    "fmc_splined_table": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, D, D, I, D, D, I]),
    "fmc_pair_probe": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, D, D, D]),
    "fmc_parse_state": (C.c_int, [C.c_char_p, I, I, D, D, D, I, I, I, I, D]),
    "fmc_load_state": (C.c_int, [C.c_void_p, C.c_char_p]),
    "fmc_server_create": (C.c_int, [C.c_int, C.c_longlong, C.c_int, C.POINTER(C.c_void_p)]),
    "fmc_server_destroy": (C.c_int, [C.c_void_p]),
    "fmc_server_set_split": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "fmc_server_submit": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int, D, D, I, C.c_int, I, I, I, C.c_longlong,
                                    C.c_ulonglong, I]),
    "fmc_server_submit_moves": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int, D, D, I, C.c_int, I, I, I, C.c_longlong,
                                          C.c_ulonglong, C.c_int, I, I, C.POINTER(C.c_uint), D, D, I]),
    "fmc_server_step": (C.c_int, [C.c_void_p, C.c_longlong, I, I]),
    "fmc_server_job": (C.c_int, [C.c_void_p, C.c_int, I, D, C.POINTER(C.c_ulonglong), C.POINTER(C.c_ulonglong),
                                 D, D, D]),
    "fmc_server_job_positions": (C.c_int, [C.c_void_p, C.c_int, D]),
    "fmc_server_job_error": (C.c_int, [C.c_void_p, C.c_int, C.c_char_p, C.c_int]),
    "fmc_server_status": (C.c_int, [C.c_void_p, D, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
}


class Change(C.Structure):
    """ctypes mirror of ``fmc_change`` (Faunus::Change)."""
    _fields_ = [("everything", C.c_int), ("volume_change", C.c_int), ("matter_change", C.c_int),
                ("num_groups", C.c_int), ("group_index", I), ("all", U8), ("internal", U8),
                ("offsets", I), ("indices", I)]


def lib():
    """Load the engine library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                              "(hipcc --offload-arch=gfx950); there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().fmc_last_error().decode())


def _p(a, t=C.c_double):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def device_count():
    n = C.c_int()
    _check(lib().fmc_device_count(C.byref(n)))
    return n.value


def copy_probe(nbytes, repeats=20):
    """ms per launch of a streaming copy of `nbytes` bytes (HBM counter calibration)."""
    ms = C.c_double()
    _check(lib().fmc_copy_probe(int(nbytes), int(repeats), C.byref(ms)))
    return ms.value


def _particles(xyz, charge, atom_id):
    """Contiguous particle arrays with matching lengths (3n positions, n charges, n ids)."""
    xyz = np.ascontiguousarray(xyz, np.float64)
    charge = np.ascontiguousarray(charge, np.float64).ravel()
    atom_id = np.ascontiguousarray(atom_id, np.int32).ravel()
    n = len(charge)
    if xyz.size != 3 * n or len(atom_id) != n:
        raise ValueError(f"particle arrays disagree: {xyz.size // 3} positions, {n} charges, {len(atom_id)} ids")
    return xyz, charge, atom_id


def _groups(begin, size, molid, mass_center=None):
    begin = np.ascontiguousarray(begin, np.int32).ravel()
    size = np.ascontiguousarray(size, np.int32).ravel()
    molid = np.ascontiguousarray(molid, np.int32).ravel()
    if not (len(begin) == len(size) == len(molid)):
        raise ValueError("group arrays disagree in length")
    cm = None if mass_center is None else np.ascontiguousarray(mass_center, np.float64)
    if cm is not None and cm.size != 3 * len(begin):
        raise ValueError("mass_center needs 3 values per group")
    return begin, size, molid, cm


def make_change(change):
    """Build an ``fmc_change`` from a dict {"everything": bool, "groups": [{group_index,
    all, internal, relative_atom_indices}]}; returns (struct, keep-alive arrays)."""
    groups = change.get("groups", [])
    gi = np.array([g["group_index"] for g in groups], np.int32)
    al = np.array([int(g.get("all", False)) for g in groups], np.uint8)
    it = np.array([int(g.get("internal", False)) for g in groups], np.uint8)
    off = np.zeros(len(groups) + 1, np.int32)
    idx = []
    for k, g in enumerate(groups):
        rel = list(g.get("relative_atom_indices", []))
        idx += rel
        off[k + 1] = off[k] + len(rel)
    idx = np.array(idx if idx else [0], np.int32)
    c = Change(int(change.get("everything", False)), int(change.get("volume_change", False)),
               int(change.get("matter_change", False)), len(groups), _p(gi, C.c_int), _p(al, C.c_ubyte),
               _p(it, C.c_ubyte), _p(off, C.c_int), _p(idx, C.c_int))
    return c, (gi, al, it, off, idx)


class Engine:
    """One Metropolis system on the GPU (both MC states)."""

    def __init__(self, config):
        text = config if isinstance(config, str) else json.dumps(config)
        h = C.c_void_p()
        _check(lib().fmc_create(text.encode(), C.byref(h)))
        self.h = h
        self.n = 0

    def close(self):
        if getattr(self, "h", None):
            lib().fmc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_particles(self, xyz, charge, atom_id):
        xyz, charge, atom_id = _particles(xyz, charge, atom_id)
        self.n = len(charge)
        _check(lib().fmc_set_particles(self.h, self.n, _p(xyz), _p(charge), _p(atom_id, C.c_int)))

    def set_groups(self, begin, size, molid, mass_center=None):
        begin, size, molid, cm = _groups(begin, size, molid, mass_center)
        _check(lib().fmc_set_groups(self.h, len(begin), _p(begin, C.c_int), _p(size, C.c_int),
                                    _p(molid, C.c_int), _p(cm)))

    def load_state(self, text):
        """Particles and groups from a reference state file (Space JSON text or dict)."""
        if not isinstance(text, str):
            text = json.dumps(text)
        _check(lib().fmc_load_state(self.h, text.encode()))
        n = C.c_int()
        _check(lib().fmc_parse_state(text.encode(), C.byref(n), None, None, None, None, None, None, None, None, None))
        self.n = n.value

    def init(self):
        e = C.c_double()
        _check(lib().fmc_init(self.h, C.byref(e)))
        return e.value

    @property
    def num_terms(self):
        n = C.c_int()
        _check(lib().fmc_num_terms(self.h, C.byref(n)))
        return n.value

    def term_names(self):
        out = []
        for i in range(self.num_terms):
            buf = C.create_string_buffer(64)
            _check(lib().fmc_term_name(self.h, i, buf, 64))
            out.append(buf.value.decode())
        return out

    def stage_trial(self, index, xyz):
        index = np.ascontiguousarray(np.atleast_1d(index), np.int32)
        xyz = np.ascontiguousarray(np.atleast_2d(xyz), np.float64)
        _check(lib().fmc_stage_trial(self.h, len(index), _p(index, C.c_int), _p(xyz)))

    def replay_moves(self, group, offsets, index, xyz, uniform, whole=None, mass_center=None):
        """Metropolis replay of multi-atom moves (see fmc_replay_moves); returns du, accepted."""
        n = len(uniform)
        group = np.ascontiguousarray(group, np.int32)
        offsets = np.ascontiguousarray(offsets, np.int64)
        index = np.ascontiguousarray(index, np.int32)
        xyz = np.ascontiguousarray(xyz, np.float64)
        uniform = np.ascontiguousarray(uniform, np.float64)
        if len(group) != n or len(offsets) != n + 1 or len(index) < offsets[-1] or xyz.size < 3 * offsets[-1]:
            raise ValueError("replay_moves: inconsistent array lengths")
        wh = None if whole is None else np.ascontiguousarray(whole, np.uint8)
        cm = None if mass_center is None else np.ascontiguousarray(mass_center, np.float64)
        du = np.zeros(n)
        acc = np.zeros(n, np.uint8)
        _check(lib().fmc_replay_moves(self.h, n, _p(group, C.c_int), _p(offsets, C.c_long), _p(index, C.c_int),
                                      _p(xyz), _p(cm), _p(wh, C.c_ubyte), _p(uniform), _p(du),
                                      _p(acc, C.c_ubyte)))
        return du, acc

    def stage_mass_center(self, group, cm):
        cm = np.ascontiguousarray(cm, np.float64)
        if cm.size != 3:
            raise ValueError("mass centre needs 3 values")
        _check(lib().fmc_stage_mass_center(self.h, int(group), _p(cm)))

    def mass_center(self, group):
        cm = np.zeros(3)
        _check(lib().fmc_get_mass_center(self.h, int(group), _p(cm)))
        return cm

    def delta_u(self, change):
        c, keep = make_change(change)
        en = np.zeros(self.num_terms)
        eo = np.zeros(self.num_terms)
        _check(lib().fmc_delta_u(self.h, C.byref(c), _p(en), _p(eo)))
        return en, eo

    def energy(self, change, state):
        c, keep = make_change(change)
        t = np.zeros(self.num_terms)
        tot = C.c_double()
        _check(lib().fmc_energy(self.h, C.byref(c), state, _p(t), C.byref(tot)))
        return tot.value, t

    def sync(self, accept):
        _check(lib().fmc_sync(self.h, int(bool(accept))))

    def total_energy(self):
        t = np.zeros(self.num_terms)
        tot = C.c_double()
        _check(lib().fmc_total_energy(self.h, _p(t), C.byref(tot)))
        return tot.value, t

    def check_q(self):
        a, b = C.c_double(), C.c_double()
        _check(lib().fmc_check_q(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def positions(self):
        xyz = np.zeros((self.n, 3))
        _check(lib().fmc_get_positions(self.h, _p(xyz)))
        return xyz

    def replay_translate(self, group, index, xyz, uniform, terms=True):
        n = len(uniform)
        nt = self.num_terms
        group = np.ascontiguousarray(group, np.int32)
        index = np.ascontiguousarray(index, np.int32)
        xyz = np.ascontiguousarray(xyz, np.float64)
        uniform = np.ascontiguousarray(uniform, np.float64)
        out = dict(du=np.zeros(n), accepted=np.zeros(n, np.uint8))
        if terms:
            out["e_new"] = np.zeros((n, nt))
            out["e_old"] = np.zeros((n, nt))
        _check(lib().fmc_replay_translate(self.h, n, _p(group, C.c_int), _p(index, C.c_int), _p(xyz), _p(uniform),
                                          _p(out["du"]), _p(out["accepted"], C.c_ubyte), _p(out.get("e_new")),
                                          _p(out.get("e_old"))))
        return out

    def run_translate(self, group, index, displacement, uniform):
        """Plain Metropolis run (one GPU evaluation per move) from displacements."""
        n = len(uniform)
        group = np.ascontiguousarray(group, np.int32)
        index = np.ascontiguousarray(index, np.int32)
        displacement = np.ascontiguousarray(displacement, np.float64)
        uniform = np.ascontiguousarray(uniform, np.float64)
        du = np.zeros(n)
        acc = np.zeros(n, np.uint8)
        _check(lib().fmc_run_translate(self.h, n, _p(group, C.c_int), _p(index, C.c_int), _p(displacement),
                                       _p(uniform), _p(du), _p(acc, C.c_ubyte)))
        return du, acc

    def set_timing(self, on):
        _check(lib().fmc_set_timing(self.h, int(bool(on))))

    def timing(self):
        ms, n = C.c_double(), C.c_long()
        _check(lib().fmc_get_timing(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def coulomb_spline(self):
        cap = 256
        k = np.zeros(cap)
        c = np.zeros(6 * cap)
        n = C.c_int()
        _check(lib().fmc_coulomb_spline(self.h, cap, _p(k), _p(c), C.byref(n)))
        return k[: n.value], c[: 6 * max(n.value - 1, 0)]

    def pair_probe(self, type_a, type_b, qa, qb, r2):
        """Device pair energies at squared distances r2: (FP64 form, FP32 form)."""
        r2 = np.ascontiguousarray(r2, np.float64)
        u64, u32 = np.zeros(len(r2)), np.zeros(len(r2))
        _check(lib().fmc_pair_probe(self.h, int(type_a), int(type_b), float(qa), float(qb), len(r2), _p(r2), _p(u64),
                                    _p(u32)))
        return u64, u32

    def splined_table(self, type_a, type_b, cap=256):
        """The engine's nonbonded_splined table of a type pair."""
        k = np.zeros(cap)
        c = np.zeros(6 * cap)
        n, hs = C.c_int(), C.c_int()
        r0, r1 = C.c_double(), C.c_double()
        _check(lib().fmc_splined_table(self.h, int(type_a), int(type_b), cap, _p(k), _p(c), C.byref(n), C.byref(r0),
                                       C.byref(r1), C.byref(hs)))
        m = n.value
        return dict(knots=k[:m], coeff=c[:6 * max(m - 1, 0)], rmin2=r0.value, rmax2=r1.value, hardsphere=bool(hs.value))

    @property
    def num_kvectors(self):
        n = C.c_int()
        _check(lib().fmc_num_kvectors(self.h, C.byref(n)))
        return n.value


class Batch:
    """Replicas advanced by the device-resident Metropolis kernel (one workgroup each)."""

    def __init__(self, fp32_pairs=False, cell_divisions=1, cell_capacity=0):
        """``cell_divisions`` 2 gives replicas with a cell list cells of half the pair
        range and 125-cell stencils (move-set replicas only); ``cell_capacity`` > 0 fixes
        the members per cell (0: automatic)."""
        h = C.c_void_p()
        _check(lib().fmc_batch_create(int(bool(fp32_pairs)), C.byref(h)))
        self.h = h
        self.sizes = []
        self.ngroups = []
        self.nterms = []
        if cell_divisions != 1:
            _check(lib().fmc_batch_set_cell_divisions(self.h, int(cell_divisions)))
        if cell_capacity:
            _check(lib().fmc_batch_set_cell_capacity(self.h, int(cell_capacity)))

    def close(self):
        if getattr(self, "h", None):
            lib().fmc_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, system, beta=1.0, seed=1, ladder=-1, moves=None):
        """Add a replica; ``moves`` (list of dicts with kind, molid, weight, dp, dprot)
        selects the general move set, otherwise single-particle translations."""
        cfg, xyz, q, ids, g = system
        text = cfg if isinstance(cfg, str) else json.dumps(cfg)
        xyz, q, ids = _particles(xyz, q, ids)
        gb, gs, gm, _ = _groups(g["begin"], g["size"], g["molid"])
        r = C.c_int()
        if moves is None:
            _check(lib().fmc_batch_add_replica(self.h, text.encode(), len(q), _p(xyz), _p(q), _p(ids, C.c_int),
                                               len(gb), _p(gb, C.c_int), _p(gs, C.c_int), _p(gm, C.c_int),
                                               float(beta), int(seed), int(ladder), C.byref(r)))
        else:
            kind = np.array([k["kind"] for k in moves], np.int32)
            molid = np.array([k["molid"] for k in moves], np.int32)
            weight = np.array([k.get("weight", 1) for k in moves], np.uint32)
            dp = np.array([float(k.get("dp", 0.0)) for k in moves])
            dprot = np.array([float(k.get("dprot", 0.0)) for k in moves])
            _check(lib().fmc_batch_add_replica_moves(
                self.h, text.encode(), len(q), _p(xyz), _p(q), _p(ids, C.c_int), len(gb), _p(gb, C.c_int),
                _p(gs, C.c_int), _p(gm, C.c_int), float(beta), int(seed), int(ladder), len(kind), _p(kind, C.c_int),
                _p(molid, C.c_int), _p(weight, C.c_uint), _p(dp), _p(dprot), C.byref(r)))
        self.sizes.append(len(q))
        self.ngroups.append(len(gb))
        return r.value

    def run(self, moves, trace=0):
        m = np.ascontiguousarray(np.broadcast_to(np.asarray(moves, np.int64), (len(self.sizes),)), np.int64)
        _check(lib().fmc_batch_run(self.h, m.ctypes.data_as(C.POINTER(C.c_longlong)), int(trace)))

    def run_split(self, moves, groups=32, stride=1, trace=0):
        _check(lib().fmc_batch_run_split(self.h, int(moves), int(groups), int(stride), int(trace)))

    def run_pipe(self, moves, groups=16, depth=2, trace=0):
        _check(lib().fmc_batch_run_pipe(self.h, int(moves), int(groups), int(depth), int(trace)))

    def sample_energy(self, every):
        """Device energy samples every ``every`` moves (0: off); clears the accumulators."""
        _check(lib().fmc_batch_sample_energy(self.h, int(every)))

    def energy_stats(self, r):
        """Replica r's device energy samples: count, sum, sum of squares, mean and
        variance (kT at T_ref)."""
        n, s1, s2 = C.c_ulonglong(), C.c_double(), C.c_double()
        _check(lib().fmc_batch_energy_stats(self.h, int(r), C.byref(n), C.byref(s1), C.byref(s2)))
        k = n.value
        mean = s1.value / k if k else float("nan")
        return dict(samples=k, sum=s1.value, sum_sq=s2.value, mean=mean,
                    variance=(s2.value / k - mean * mean) if k else float("nan"))

    def rdf(self, r, type_a, type_b, dr, nbins, counts=None):
        """Add the a-b pair-distance histogram of replica r's current configuration to
        ``counts`` (uint64, nbins; a new array if None) and return it."""
        if counts is None:
            counts = np.zeros(int(nbins), np.uint64)
        if counts.dtype != np.uint64 or len(counts) != int(nbins) or not counts.flags["C_CONTIGUOUS"]:
            raise ValueError("counts must be a contiguous uint64 array of nbins")
        _check(lib().fmc_batch_rdf(self.h, int(r), int(type_a), int(type_b), float(dr), int(nbins),
                                   counts.ctypes.data_as(C.POINTER(C.c_ulonglong))))
        return counts

    def split_error(self):
        e = C.c_int()
        _check(lib().fmc_batch_split_error(self.h, C.byref(e)))
        return e.value

    def sync(self):
        ms = C.c_double()
        _check(lib().fmc_batch_sync(self.h, C.byref(ms)))
        return ms.value

    def set_timing(self, on):
        _check(lib().fmc_batch_set_timing(self.h, int(bool(on))))

    def capture_interval(self, moves, seed, round0=0):
        """Capture one replica-exchange interval (moves per replica, then one round of
        ladder swaps with a device round counter) as a HIP graph."""
        m = np.ascontiguousarray(np.broadcast_to(np.asarray(moves, np.int64), (len(self.sizes),)), np.int64)
        _check(lib().fmc_batch_capture_interval(self.h, m.ctypes.data_as(C.POINTER(C.c_longlong)), int(seed),
                                                int(round0)))

    def replay(self, count=1):
        _check(lib().fmc_batch_replay(self.h, int(count)))

    def run_ladders(self, moves, intervals, seed, round0=0):
        """Free-running ladders: replica r runs intervals[r] intervals of moves[r] moves in
        one launch, each ladder swapping its rungs after each of its own intervals (round
        round0 + interval).  Blocks until done."""
        n = len(self.sizes)
        m = np.ascontiguousarray(np.broadcast_to(np.asarray(moves, np.int64), (n,)), np.int64)
        iv = np.ascontiguousarray(np.broadcast_to(np.asarray(intervals, np.int32), (n,)), np.int32)
        _check(lib().fmc_batch_run_ladders(self.h, m.ctypes.data_as(C.POINTER(C.c_longlong)),
                                           iv.ctypes.data_as(C.POINTER(C.c_int)), int(seed), int(round0)))

    def swap_counts(self):
        a, b = C.c_longlong(), C.c_longlong()
        _check(lib().fmc_batch_swap_counts(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def autotune(self, trial_moves=200):
        """Time the candidate launch shapes and keep the fastest: returns the chosen value
        and {candidate: us per replica-move} (empty when there is no choice)."""
        cap = 8
        n, chosen = C.c_int(), C.c_int()
        vals = np.zeros(cap, np.int32)
        us = np.zeros(cap)
        _check(lib().fmc_batch_autotune(self.h, int(trial_moves), cap, C.byref(n), _p(vals, C.c_int), _p(us),
                                        C.byref(chosen)))
        return chosen.value, {int(vals[i]): float(us[i]) for i in range(min(n.value, cap))}

    def swap(self, seed, rnd):
        a, b = C.c_longlong(), C.c_longlong()
        _check(lib().fmc_batch_swap(self.h, int(seed), int(rnd), C.byref(a), C.byref(b)))
        return a.value, b.value

    def info(self, r):
        e, b = C.c_double(), C.c_double()
        m, a = C.c_ulonglong(), C.c_ulonglong()
        err = C.c_int()
        _check(lib().fmc_batch_replica_info(self.h, r, C.byref(e), C.byref(b), C.byref(m), C.byref(a), C.byref(err)))
        return dict(energy=e.value, beta=b.value, moves=m.value, accepted=a.value, error=err.value)

    def positions(self, r):
        xyz = np.zeros((self.sizes[r], 3))
        _check(lib().fmc_batch_positions(self.h, r, _p(xyz)))
        return xyz

    def get_state(self, r):
        """Checkpoint of replica r: positions, mass centres, move counter, accepted count."""
        xyz = np.zeros((self.sizes[r], 3))
        cm = np.zeros((self.ngroups[r], 3))
        m, a = C.c_ulonglong(), C.c_ulonglong()
        _check(lib().fmc_batch_get_replica_state(self.h, int(r), _p(xyz), _p(cm), C.byref(m), C.byref(a)))
        return dict(xyz=xyz, mass_center=cm, moves=m.value, accepted=a.value)

    def set_state(self, r, state):
        """Resume replica r from a checkpoint made by get_state (mass_center may be None)."""
        xyz = np.ascontiguousarray(state["xyz"], np.float64)
        cm = state.get("mass_center")
        cm = None if cm is None else np.ascontiguousarray(cm, np.float64)
        _check(lib().fmc_batch_set_replica_state(self.h, int(r), _p(xyz), None if cm is None else _p(cm),
                                                 int(state["moves"]), int(state.get("accepted", 0))))

    def trace(self, r, count):
        idx = np.zeros(count, np.int32)
        acc = np.zeros(count, np.uint8)
        du = np.zeros(count)
        xyz = np.zeros((count, 3))
        _check(lib().fmc_batch_trace(self.h, r, count, _p(idx, C.c_int), _p(acc, C.c_ubyte), _p(du), _p(xyz)))
        return dict(index=idx, accepted=acc, du=du, xyz=xyz)

    def total_energy(self, r, nterms=3):
        t = np.zeros(8)
        tot = C.c_double()
        _check(lib().fmc_batch_total_energy(self.h, r, _p(t), C.byref(tot)))
        return tot.value, t

    def check_q(self, r):
        a, b = C.c_double(), C.c_double()
        _check(lib().fmc_batch_check_q(self.h, r, C.byref(a), C.byref(b)))
        return a.value, b.value


def parse_state(text):
    """Read a reference state file (Space JSON): box, particles and groups as arrays."""
    if not isinstance(text, str):
        text = json.dumps(text)
    n, ng = C.c_int(), C.c_int()
    _check(lib().fmc_parse_state(text.encode(), C.byref(n), C.byref(ng), None, None, None, None, None, None, None,
                                 None))
    box = np.zeros(3)
    xyz = np.zeros((n.value, 3))
    q = np.zeros(n.value)
    ids = np.zeros(n.value, np.int32)
    gb, gs, gm = (np.zeros(ng.value, np.int32) for _ in range(3))
    cm = np.zeros((ng.value, 3))
    _check(lib().fmc_parse_state(text.encode(), C.byref(n), C.byref(ng), _p(box), _p(xyz), _p(q), _p(ids, C.c_int),
                                 _p(gb, C.c_int), _p(gs, C.c_int), _p(gm, C.c_int), _p(cm)))
    return dict(box=box, xyz=xyz, q=q, ids=ids, groups=dict(begin=gb, size=gs, molid=gm), mass_center=cm)


class Server:
    """Continuous-batching serving loop: jobs are translational Metropolis runs that
    share persistent launches while their HBM footprint fits the budget."""

    QUEUED, RESIDENT, DONE, FAILED = 0, 1, 2, 3

    def __init__(self, fp32_pairs=True, hbm_budget_bytes=64 << 30, max_slots=1024, split_atoms=0, max_parts=16):
        """``split_atoms`` > 0 shares each job with a general move set and no cell list
        among ceil(n / split_atoms) workgroups (at most ``max_parts``)."""
        h = C.c_void_p()
        _check(lib().fmc_server_create(int(bool(fp32_pairs)), int(hbm_budget_bytes), int(max_slots), C.byref(h)))
        self.h = h
        self.sizes = {}
        if split_atoms:
            _check(lib().fmc_server_set_split(self.h, int(split_atoms), int(max_parts)))

    def close(self):
        if getattr(self, "h", None):
            lib().fmc_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, system, moves, seed, kinds=None):
        """Queue a job of ``moves`` trial moves; ``kinds`` (list of dicts with kind,
        molid, weight, dp, dprot) selects the general move set."""
        cfg, xyz, q, ids, g = system
        text = cfg if isinstance(cfg, str) else json.dumps(cfg)
        xyz, q, ids = _particles(xyz, q, ids)
        gb, gs, gm, _ = _groups(g["begin"], g["size"], g["molid"])
        j = C.c_int()
        if kinds is None:
            _check(lib().fmc_server_submit(self.h, text.encode(), len(q), _p(xyz), _p(q), _p(ids, C.c_int), len(gb),
                                           _p(gb, C.c_int), _p(gs, C.c_int), _p(gm, C.c_int), int(moves), int(seed),
                                           C.byref(j)))
        else:
            kind = np.array([k["kind"] for k in kinds], np.int32)
            molid = np.array([k["molid"] for k in kinds], np.int32)
            weight = np.array([k.get("weight", 1) for k in kinds], np.uint32)
            dp = np.array([float(k.get("dp", 0.0)) for k in kinds])
            dprot = np.array([float(k.get("dprot", 0.0)) for k in kinds])
            _check(lib().fmc_server_submit_moves(
                self.h, text.encode(), len(q), _p(xyz), _p(q), _p(ids, C.c_int), len(gb), _p(gb, C.c_int),
                _p(gs, C.c_int), _p(gm, C.c_int), int(moves), int(seed), len(kind), _p(kind, C.c_int),
                _p(molid, C.c_int), _p(weight, C.c_uint), _p(dp), _p(dprot), C.byref(j)))
        self.sizes[j.value] = len(q)
        return j.value

    def step(self, quantum=2000):
        r, q = C.c_int(), C.c_int()
        _check(lib().fmc_server_step(self.h, int(quantum), C.byref(r), C.byref(q)))
        return r.value, q.value

    def job(self, j):
        st, e = C.c_int(), C.c_double()
        m, a = C.c_ulonglong(), C.c_ulonglong()
        t0, t1, t2 = C.c_double(), C.c_double(), C.c_double()
        _check(lib().fmc_server_job(self.h, int(j), C.byref(st), C.byref(e), C.byref(m), C.byref(a), C.byref(t0),
                                    C.byref(t1), C.byref(t2)))
        out = dict(state=st.value, energy=e.value, moves=m.value, accepted=a.value, t_submit=t0.value,
                   t_admit=t1.value, t_done=t2.value)
        if st.value == self.FAILED:
            buf = C.create_string_buffer(512)
            _check(lib().fmc_server_job_error(self.h, int(j), buf, 512))
            out["error"] = buf.value.decode()
        return out

    def positions(self, j):
        xyz = np.zeros((self.sizes[j], 3))
        _check(lib().fmc_server_job_positions(self.h, int(j), _p(xyz)))
        return xyz

    def status(self):
        t = C.c_double()
        m, b = C.c_longlong(), C.c_longlong()
        _check(lib().fmc_server_status(self.h, C.byref(t), C.byref(m), C.byref(b)))
        return dict(seconds=t.value, moves_done=m.value, bytes_in_use=b.value)
