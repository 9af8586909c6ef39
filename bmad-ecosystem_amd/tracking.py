"""Host API over libbmx.so, named after the Bmad calls it replaces.

  Context          one GPU + one HIP stream (bmx_ctx)
  Lattice          flattened element table of lat%branch(ix)%ele(1:n) (bmx_lattice)
  Beam             SoA particle state resident in HBM (bmx_beam)
  track_beam       n_turns full turns, persistent kernel
                   (bmad/multiparticle/beam_mod.f90:40 track_beam, per turn as in
                   bsim/modules/lt_tracking_mod.f90:1631-1716)
  track_beam_from  the same with every turn starting after a given element (LTT's ele_start)
  track_bunch      one element span ele1 -> ele2 (beam_mod.f90:103), wrapping round the ring
  track1_bunch     one element for the whole bunch (beam_mod.f90:238 -> beam_utils.f90:32)
  track_beam_plain n_turns as n_turns x n_ele element launches (the reference loop nest)
  track_all        single particles through the lattice, stopping at loss
                   (bmad/code/track_all.f90:25-145)

Loss is data: ``state`` holds Bmad's coord_struct%state codes, ``loss_ele`` the 1-based
element index and ``loss_turn`` the 0-based turn (lt_tracking_mod.f90:1395).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import elements as E
from . import native

FLOAT_KEYS = ("x", "px", "y", "py", "z", "pz", "beta")
INT_KEYS = ("state", "loss_ele", "loss_turn")
KEYS = FLOAT_KEYS + INT_KEYS


class Context:
    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        native.call("bmx_ctx_create", device, ctypes.byref(h))
        self.h = h
        self.device = device
        self.maps = "fast"

    def synchronize(self):
        native.call("bmx_ctx_synchronize", self.h)

    @property
    def stream(self) -> int:
        return native.load().bmx_ctx_stream(self.h) or 0

    def set_maps(self, form: str) -> None:
        """'fast' (default) or 'reference' formulation of the element maps."""
        native.call("bmx_ctx_set_maps", self.h, {"fast": 1, "reference": 2}[form])
        self.maps = form

    def set_run(self, chunk_turns: int = 5, grid: int = 0, waves: int = 0) -> None:
        """Turns per chunk of bmx_track_turns (0 = one chunk), persistent grid size (0 = device
        occupancy limit) and waves per SIMD of the kernel variant (3, 4 or 5; 0 = chosen per
        call: on a specialised lattice 5 or 4 by the live particles' count, else 5 with the fast
        maps, 3 with the reference maps)."""
        native.call("bmx_ctx_set_run", self.h, int(chunk_turns), int(grid), int(waves))

    def last_run(self):
        """(turns per chunk, workgroups, waves per SIMD) of the last persistent tracking call."""
        c, g, w = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        native.call("bmx_ctx_last_run", self.h, ctypes.byref(c), ctypes.byref(g), ctypes.byref(w))
        return c.value, g.value, w.value

    def profile(self, on: bool = True) -> None:
        """Bracket every tracking-kernel launch of track_beam with HIP events."""
        native.call("bmx_ctx_profile", self.h, 1 if on else 0)

    def kernel_time(self):
        """(total ms, launches) of the tracking kernel since the last call (synchronises)."""
        ms, n = ctypes.c_double(), ctypes.c_int64()
        native.call("bmx_ctx_kernel_time", self.h, ctypes.byref(ms), ctypes.byref(n))
        return ms.value, n.value

    def record(self, slot: int) -> None:
        """Record HIP event `slot` on this context's stream."""
        native.call("bmx_ctx_record", self.h, int(slot))

    def elapsed_ms(self, slot0: int, slot1: int) -> float:
        ms = ctypes.c_float()
        native.call("bmx_ctx_elapsed_ms", self.h, int(slot0), int(slot1), ctypes.byref(ms))
        return ms.value

    def take_work(self) -> int:
        """Particle-element passes tracked since the last call (synchronises)."""
        v = ctypes.c_int64()
        native.call("bmx_ctx_take_work", self.h, ctypes.byref(v))
        return v.value

    def close(self):
        if self.h:
            native.call("bmx_ctx_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rad_config(damping: bool = True, fluctuations: bool = False, zero_average: bool = False,
               scale: float = 1.0, seed: int = 0) -> native.RadConfig:
    """bmx_rad_config: bmad_com%radiation_damping_on, %radiation_fluctuations_on,
    %radiation_zero_average, %synch_rad_scale and the fluctuation stream's seed."""
    return native.RadConfig(int(bool(damping)), int(bool(fluctuations)), int(bool(zero_average)), 0,
                            float(scale), int(seed) & (2**64 - 1))


def rad_maps(n_ele: int) -> np.ndarray:
    """Zeroed radiation maps for n_ele elements, shape (n_ele, 2): [i, 0] = rm0 (upstream half),
    [i, 1] = rm1 (downstream half) of element i, fields as rad_map_struct."""
    return np.zeros((n_ele, 2), dtype=native.RAD_MAP_DTYPE)


class Lattice:
    """A lattice on the device (bmx_lattice_create). With radiation maps (rad_maps(), filled
    from each element's rad_map%rm0 / %rm1) and settings (rad_config()), the elements that radiate
    take track1_radiation's kicks at both ends (bmx_lattice_create_rad)."""

    def __init__(self, ctx: Context, eles, mc2: float = E.M_ELECTRON, charge: int = -1,
                 rad_maps: np.ndarray | None = None, rad: native.RadConfig | None = None):
        arr = eles if isinstance(eles, np.ndarray) else E.to_array(eles)
        self.arr = np.ascontiguousarray(arr)
        self.ctx = ctx
        h = ctypes.c_void_p()
        if rad_maps is None:
            native.call("bmx_lattice_create", ctx.h, self.arr.ctypes.data, len(self.arr), mc2,
                        charge, ctypes.byref(h))
        else:
            maps = np.ascontiguousarray(rad_maps, dtype=native.RAD_MAP_DTYPE)
            if maps.shape != (len(self.arr), 2):
                raise ValueError(f"rad_maps must have shape ({len(self.arr)}, 2), not {maps.shape}")
            cfg = rad if rad is not None else rad_config()
            native.call("bmx_lattice_create_rad", ctx.h, self.arr.ctypes.data, len(self.arr), mc2,
                        charge, maps.ctypes.data, ctypes.byref(cfg), ctypes.byref(h))
        self.h = h
        n = ctypes.c_int32()
        mask = ctypes.c_uint64()
        native.call("bmx_lattice_info", self.h, ctypes.byref(n), ctypes.byref(mask))
        self.n_ele = n.value
        self.kind_mask = mask.value
        self.specialized = False

    def set_radiation(self, rad: native.RadConfig) -> None:
        """New radiation settings for later tracking calls (bmx_lattice_set_radiation)."""
        native.call("bmx_lattice_set_radiation", self.h, ctypes.byref(rad))

    def specialize(self) -> bool:
        """Compile tile kernels specialised to this lattice's repeated cell (bmx_lattice_specialize).

        Returns False when the lattice has no cell repeated over three quarters of it; the
        generic kernel is then used. Other errors raise."""
        rc = native.load().bmx_lattice_specialize(self.h)
        if rc == native.ERR_UNSUPPORTED:
            return False
        native.check(rc)
        self.specialized = True
        return True

    def close(self):
        if self.h:
            native.call("bmx_lattice_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _host_view(beam: dict) -> native.BeamView:
    for k in FLOAT_KEYS:
        assert beam[k].dtype == np.float64 and beam[k].flags.c_contiguous, k
    for k in INT_KEYS:
        assert beam[k].dtype == np.int32 and beam[k].flags.c_contiguous, k
    return native.BeamView(*[beam[k].ctypes.data for k in KEYS])


class Beam:
    """Particle state in HBM. Either library-owned (n particles) or wrapping device tensors."""

    def __init__(self, ctx: Context, n: int | None = None, tensors: dict | None = None):
        self.ctx = ctx
        self.tensors = tensors
        h = ctypes.c_void_p()
        if tensors is not None:
            n = int(tensors["x"].shape[0])
            view = native.BeamView(*[int(tensors[k].data_ptr()) for k in KEYS])
            native.call("bmx_beam_wrap", ctx.h, ctypes.byref(view), n, ctypes.byref(h))
        else:
            native.call("bmx_beam_create", ctx.h, int(n), ctypes.byref(h))
        self.h = h
        self.n = int(n)

    @classmethod
    def from_host(cls, ctx: Context, beam: dict) -> "Beam":
        b = cls(ctx, int(beam["x"].shape[0]))
        b.upload(beam)
        return b

    def upload(self, beam: dict):
        v = _host_view(beam)
        native.call("bmx_beam_upload", self.h, ctypes.byref(v), int(beam["x"].shape[0]))

    def download(self, n: int | None = None) -> dict:
        n = self.n if n is None else n
        out = {k: np.empty(n, dtype=np.float64) for k in FLOAT_KEYS}
        out.update({k: np.empty(n, dtype=np.int32) for k in INT_KEYS})
        v = _host_view(out)
        native.call("bmx_beam_download", self.h, ctypes.byref(v), n)
        return out

    def count_alive(self) -> int:
        c = ctypes.c_int64()
        native.call("bmx_beam_count_alive", self.h, ctypes.byref(c))
        return c.value

    def close(self):
        if self.h:
            native.call("bmx_beam_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def track_beam(ctx: Context, lat: Lattice, beam: Beam, n_turns: int, turn0: int = 0) -> None:
    """Persistent tracking of n_turns full turns (asynchronous on ctx's stream)."""
    native.call("bmx_track_turns", ctx.h, lat.h, beam.h, int(n_turns), int(turn0))


def track_beam_from(ctx: Context, lat: Lattice, beam: Beam, ix_ele_start: int, n_turns: int,
                    turn0: int = 0) -> None:
    """n_turns turns that each start after element ix_ele_start: LTT's
    track_bunch(lat, bunch, ele_start, ele_start) per turn (bmx_track_turns_from; asynchronous)."""
    native.call("bmx_track_turns_from", ctx.h, lat.h, beam.h, int(ix_ele_start), int(n_turns),
                int(turn0))


def track_bunch(ctx: Context, lat: Lattice, beam: Beam, ix_ele1: int = 0, ix_ele2: int | None = None,
                turn: int = 0) -> None:
    """track_bunch (beam_mod.f90:103): from the end of element ix_ele1 (not tracked) to the end
    of ix_ele2 (default: the last element), wrapping round the ring when ix_ele2 <= ix_ele1
    (bmx_track_bunch; asynchronous)."""
    e2 = lat.n_ele if ix_ele2 is None else int(ix_ele2)
    native.call("bmx_track_bunch", ctx.h, lat.h, beam.h, int(ix_ele1), e2, int(turn))


def track1_bunch(ctx: Context, lat: Lattice, beam: Beam, ix_ele: int, turn: int = 0) -> None:
    """One element (1-based ix_ele) for the whole bunch (asynchronous)."""
    native.call("bmx_track_element", ctx.h, lat.h, beam.h, int(ix_ele), int(turn))


def track_beam_plain(ctx: Context, lat: Lattice, beam: Beam, n_turns: int, turn0: int = 0) -> None:
    """n_turns as element-by-element launches, state through HBM (asynchronous)."""
    native.call("bmx_track_turns_plain", ctx.h, lat.h, beam.h, int(n_turns), int(turn0))


def read_beam_file(path: str, p0c: float, mc2: float = E.M_ELECTRON, ix_bunch: int = 1):
    """Bmad ASCII beam file (bmx_beam_file_read) -> (host SoA dict, charge_tot)."""
    nb, npart = ctypes.c_int32(), ctypes.c_int64()
    native.call("bmx_beam_file_info", path.encode(), ctypes.byref(nb), ctypes.byref(npart))
    cap = int(npart.value) if ix_bunch == 1 else 1 << 26
    out = {k: np.zeros(cap, dtype=np.float64) for k in FLOAT_KEYS}
    out.update({k: np.zeros(cap, dtype=np.int32) for k in INT_KEYS})
    v = _host_view(out)
    n, q = ctypes.c_int64(), ctypes.c_double()
    native.call("bmx_beam_file_read", path.encode(), int(ix_bunch), ctypes.byref(v), cap, float(p0c),
                float(mc2), ctypes.byref(n), ctypes.byref(q))
    return {k: a[:n.value].copy() for k, a in out.items()}, q.value


def write_beam_file(path: str, beam: dict, p0c: float, species: str = "Electron",
                    charge: float = 0.0, ix_ele: int = 0, append: bool = False) -> None:
    """Host SoA dict -> Bmad ASCII column-layout beam file (bmx_beam_file_write)."""
    v = _host_view(beam)
    native.call("bmx_beam_file_write", path.encode(), ctypes.byref(v), int(beam["x"].shape[0]),
                species.encode(), float(charge), float(p0c), int(ix_ele), int(bool(append)))


def read_beam_h5(path: str, mc2: float = E.M_ELECTRON, ix_bunch: int = 1):
    """openPMD HDF5 beam file (bmx_beam_h5_read) -> (host SoA dict, p0c, charge_tot, species)."""
    nb, npart = ctypes.c_int32(), ctypes.c_int64()
    sp = ctypes.create_string_buffer(64)
    native.call("bmx_beam_h5_info", path.encode(), ctypes.byref(nb), ctypes.byref(npart), sp, 64)
    cap = int(npart.value) if ix_bunch == 1 else 1 << 26
    out = {k: np.zeros(cap, dtype=np.float64) for k in FLOAT_KEYS}
    out.update({k: np.zeros(cap, dtype=np.int32) for k in INT_KEYS})
    v = _host_view(out)
    n, p0c, q = ctypes.c_int64(), ctypes.c_double(), ctypes.c_double()
    native.call("bmx_beam_h5_read", path.encode(), int(ix_bunch), ctypes.byref(v), cap, float(mc2),
                ctypes.byref(n), ctypes.byref(p0c), ctypes.byref(q))
    return {k: a[:n.value].copy() for k, a in out.items()}, p0c.value, q.value, sp.value.decode()


def write_beam_h5(path: str, beam: dict, p0c: float, species: str = "electron",
                  charge: float = 0.0, ix_ele: int = 0, append: bool = False) -> None:
    """Host SoA dict -> openPMD HDF5 beam file, one bunch (bmx_beam_h5_write)."""
    v = _host_view(beam)
    native.call("bmx_beam_h5_write", path.encode(), ctypes.byref(v), int(beam["x"].shape[0]),
                species.encode(), float(charge), float(p0c), int(ix_ele), int(bool(append)))


def closed_orbit(ctx: Context, lats, n_dim: int = 4, guess=None, max_iter: int = 20,
                 rel_tol: float = 1e-12, abs_tol: float = 1e-15):
    """Closed orbits and one-turn matrices of several lattices at once (bmx_closed_orbit;
    bmad/code/closed_orbit_calc.f90:192). guess: (n, 6) starting orbits (zeros by default; with
    n_dim = 4 its pz is the energy the orbit is closed at). Returns (orbit (n, 6), one-turn matrix
    (n, 6, 6), status (n,): 0 converged, 1 not converged, 2 probe lost, 3 singular)."""
    lats = list(lats)
    n = len(lats)
    orb = np.zeros((n, 6)) if guess is None else np.array(guess, dtype=np.float64).reshape(n, 6)
    orb = np.ascontiguousarray(orb)
    mat = np.zeros((n, 6, 6))
    status = np.zeros(n, dtype=np.int32)
    handles = (ctypes.c_void_p * max(n, 1))(*[lt.h.value for lt in lats])
    dp = ctypes.POINTER(ctypes.c_double)
    native.call("bmx_closed_orbit", ctx.h, handles, n, int(n_dim), orb.ctypes.data_as(dp),
                mat.ctypes.data_as(dp), int(max_iter), float(rel_tol), float(abs_tol),
                status.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    return orb, mat, status


def normal_mode_tunes(mat, n_dim: int = 6):
    """Normal-mode tunes of a 6x6 one-turn matrix in radians, [0, 2 pi), ordered by plane
    dominance (bmx_normal_mode_tunes; make_N of mode3_mod.f90). n_dim = 4: the two transverse
    modes of its x-y block. Host only."""
    m = np.ascontiguousarray(np.asarray(mat, dtype=np.float64).reshape(6, 6))
    out = np.zeros(n_dim // 2)
    dp = ctypes.POINTER(ctypes.c_double)
    native.call("bmx_normal_mode_tunes", m.ctypes.data_as(dp), int(n_dim), out.ctypes.data_as(dp))
    return out


def chromaticity(ctx: Context, lats, delta_e: float = 1e-4, pz: float = 0.0):
    """Chromaticities and fractional tunes of several lattices at once (bmx_chromaticity;
    chrom_calc.f90). Returns (chrom (n, 2), tune (n, 2), status (n,)): a and b modes; status 0
    ok, 1..3 a closed orbit failed (bmx_closed_orbit's codes), 4 unstable."""
    lats = list(lats)
    n = len(lats)
    chrom = np.zeros((n, 2))
    tune = np.zeros((n, 2))
    status = np.zeros(n, dtype=np.int32)
    handles = (ctypes.c_void_p * max(n, 1))(*[lt.h.value for lt in lats])
    dp = ctypes.POINTER(ctypes.c_double)
    native.call("bmx_chromaticity", ctx.h, handles, n, float(delta_e), float(pz),
                chrom.ctypes.data_as(dp), tune.ctypes.data_as(dp),
                status.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    return chrom, tune, status


def track_beam_host(eles, beam: dict, n_turns: int, device: int = 0, turn0: int = 0,
                    mc2: float = E.M_ELECTRON, charge: int = -1, maps: str = "fast",
                    specialize: bool = False, rad_maps: np.ndarray | None = None,
                    rad: native.RadConfig | None = None) -> dict:
    """Convenience: host SoA dict in, host SoA dict out (tracks a copy). specialize: use the
    lattice's cell-specialised kernels when it has a repeated cell; rad_maps, rad: radiation
    kicks (Lattice)."""
    ctx = Context(device)
    ctx.set_maps(maps)
    lat = Lattice(ctx, eles, mc2, charge, rad_maps=rad_maps, rad=rad)
    if specialize:
        lat.specialize()
    b = Beam.from_host(ctx, beam)
    track_beam(ctx, lat, b, n_turns, turn0)
    ctx.synchronize()
    out = b.download()
    b.close()
    lat.close()
    ctx.close()
    return out


class Batch:
    """Many (lattice, beam range) groups tracked together in one captured launch sequence
    (bmx_batch_*). groups: list of (Lattice, Beam, first, count)."""

    def __init__(self, ctx: Context, groups):
        self.ctx = ctx
        self._keep = [(g[0], g[1]) for g in groups]
        specs = (native.GroupSpec * max(len(groups), 1))()
        for i, (lat, beam, first, count) in enumerate(groups):
            specs[i] = native.GroupSpec(lat.h.value, beam.h.value, int(first), int(count))
        h = ctypes.c_void_p()
        native.call("bmx_batch_create", ctx.h, specs, len(groups), ctypes.byref(h))
        self.h = h

    def track(self, n_turns: int, turn0: int = 0) -> None:
        native.call("bmx_batch_track", self.h, int(n_turns), int(turn0))

    def set_run(self, chunk_turns: int = 5, grid: int = 0, waves: int = 0) -> None:
        native.call("bmx_batch_set_run", self.h, int(chunk_turns), int(grid), int(waves))

    def tune(self, sample_turns: int = 20) -> dict:
        """Autotune the run shape on a scratch copy; returns the chosen shape."""
        c, g, w = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        ms = ctypes.c_float()
        native.call("bmx_batch_tune", self.h, int(sample_turns), ctypes.byref(c), ctypes.byref(g),
                    ctypes.byref(w), ctypes.byref(ms))
        return {"chunk_turns": c.value, "grid": g.value, "waves": w.value, "ms": ms.value}

    def close(self):
        if self.h:
            native.call("bmx_batch_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching tracking server (bmx_server_*): requests are admitted by device pool
    space, stay resident between scheduling chunks and advance together."""

    def __init__(self, ctx: Context, hbm_budget: int, chunk_turns: int = 100,
                 compact_turns: int = 10):
        cfg = native.ServerConfig(int(hbm_budget), int(chunk_turns), int(compact_turns))
        h = ctypes.c_void_p()
        native.call("bmx_server_create", ctx.h, ctypes.byref(cfg), ctypes.byref(h))
        self.h = h
        self.ctx = ctx
        self._held = {}

    def submit(self, lat: Lattice, beam: dict, n_turns: int, turn0: int = 0) -> int:
        """Queue a request; its input and output arrays are kept alive until wait()."""
        n = int(beam["x"].shape[0])
        inp = {k: np.ascontiguousarray(v) for k, v in beam.items()}
        out = {k: np.empty(n, dtype=np.float64) for k in FLOAT_KEYS}
        out.update({k: np.empty(n, dtype=np.int32) for k in INT_KEYS})
        vin, vout = _host_view(inp), _host_view(out)
        rid = ctypes.c_int64()
        native.call("bmx_server_submit", self.h, lat.h, ctypes.byref(vin), ctypes.byref(vout), n,
                    int(n_turns), int(turn0), ctypes.byref(rid))
        self._held[rid.value] = (inp, out, vin, vout, lat)
        return rid.value

    def poll(self, rid: int) -> bool:
        d = ctypes.c_int32()
        native.call("bmx_server_poll", self.h, int(rid), ctypes.byref(d))
        return bool(d.value)

    def wait(self, rid: int):
        """Block until done; returns (result beam dict, stats dict with times in seconds)."""
        st = native.RequestStats()
        try:
            native.call("bmx_server_wait", self.h, int(rid), ctypes.byref(st))
        finally:
            held = self._held.pop(rid, None)
        stats = {"submit_s": st.submit_ns * 1e-9, "admit_s": st.admit_ns * 1e-9,
                 "done_s": st.done_ns * 1e-9, "latency_s": (st.done_ns - st.submit_ns) * 1e-9,
                 "chunks": st.chunks}
        return held[1], stats

    def totals(self) -> dict:
        p, c, b = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        native.call("bmx_server_totals", self.h, ctypes.byref(p), ctypes.byref(c), ctypes.byref(b))
        return {"passes": p.value, "chunks": c.value, "busy_s": b.value * 1e-9}

    def close(self):
        if self.h:
            native.call("bmx_server_destroy", self.h)
            self.h = None
            self._held.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def dynamic_aperture(ctx: Context, lattices, grid_beam: dict, n_turns: int):
    """Survival turns of a dynamic-aperture grid on each lattice (one batch, one group per
    lattice): the loss turn of each lost particle (= completed turns, the 0-based turn it was
    lost in, lt_tracking_mod.f90:1395) or n_turns for survivors. Returns (survival_turn,
    loss_ele, final beam dicts) with one row per lattice."""
    n = int(grid_beam["x"].shape[0])
    beams = [Beam.from_host(ctx, grid_beam) for _ in lattices]
    batch = Batch(ctx, [(lat, b, 0, n) for lat, b in zip(lattices, beams)])
    batch.track(n_turns, 0)
    ctx.synchronize()
    outs = [b.download() for b in beams]
    batch.close()
    for b in beams:
        b.close()
    surv = np.stack([np.where(o["state"] == E.STATE_ALIVE, n_turns, o["loss_turn"]) for o in outs])
    loss_ele = np.stack([o["loss_ele"] for o in outs])
    return surv.astype(np.int32), loss_ele, outs


def track_orbit(ctx: Context, lat: Lattice, beam: Beam, ix_start: int = 1,
                ix_end: int | None = None, turn: int = 0) -> np.ndarray:
    """Elements ix_start..ix_end (1-based, inclusive) once, recording every particle after every
    element (track_many, track_many.f90:50-157). Returns an array (n, n_range, 8) of x, px, y,
    py, z, pz, beta, state; rows after the element a particle is lost in are zero. Blocking."""
    if ix_end is None:
        ix_end = lat.n_ele
    n_range = ix_end - ix_start + 1
    rec = np.zeros((beam.n, max(n_range, 0), 8), dtype=np.float64)
    native.call("bmx_track_orbit", ctx.h, lat.h, beam.h, int(ix_start), int(ix_end), int(turn),
                rec.ctypes.data_as(ctypes.c_void_p))
    return rec


def track_all(ctx: Context, lat: Lattice, vec, beta):
    """track_all (track_all.f90:25-145) for single particles (rows of vec): one pass over all
    elements. Returns (orbit, track_state): orbit[i, j] = (x, px, y, py, z, pz, beta, state) of
    particle i at the exit of element j + 1 (zero after its loss), track_state[i] = the 1-based
    element index where particle i was lost, or -1 (moving_forward$) if it was not."""
    vec = np.atleast_2d(np.asarray(vec, dtype=np.float64))
    n = vec.shape[0]
    beam = {k: np.ascontiguousarray(vec[:, i]) for i, k in enumerate(FLOAT_KEYS[:6])}
    beam["beta"] = np.ascontiguousarray(np.broadcast_to(np.asarray(beta, dtype=np.float64), (n,)))
    beam["state"] = np.full(n, E.STATE_ALIVE, dtype=np.int32)
    beam["loss_ele"] = np.zeros(n, dtype=np.int32)
    beam["loss_turn"] = np.full(n, -1, dtype=np.int32)
    b = Beam.from_host(ctx, beam)
    orbit = track_orbit(ctx, lat, b)
    out = b.download()
    b.close()
    track_state = np.where(out["state"] == E.STATE_ALIVE, -1, out["loss_ele"]).astype(np.int32)
    return orbit, track_state
