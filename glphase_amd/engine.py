"""Device-resident MH engine: inputs uploaded once into HBM, then the plain
path (one launch per MH iteration) or the batched persistent kernel run on a
HIP stream.  torch provides device memory and streams; all compute is the
HIP code of libglphase_amd.so.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import (GlpKnnDesc, GlpMhDesc, GlpWindow, GlpWorkDesc, GlpWorkWindow, c_f32p,
                   c_u32p, c_u64p, check, lib, ptr)


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise _lib.GlpError("no GPU visible: the glphase_amd engine runs only on a gfx950 device")
    return torch


def _dev(torch, a, device):
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint64:
        a = a.view(np.int64)
    elif a.dtype == np.uint32:
        a = a.view(np.int32)
    return torch.from_numpy(a).to(device)


class Job:
    """MH update of N samples over W windows.

    windows: list of dicts with 'panel' [H][wn] uint64, 'packed'
    [ceil(M*bits/32)][N] uint32 and 'tran' [M][3] float32 (host arrays).
    """

    def __init__(self, windows, n_samples, n_haps, codebook, pc, bits=8, num_cycles=1,
                 prop_mode=_lib.PROP_UNIFORM, precision=_lib.PREC_F32, klist=None, extra=None,
                 init_parents=None, trace=False, k0=0, kickstart=False, device="cuda:0"):
        torch = _torch()
        self.torch = torch
        self.device = torch.device(device)
        self.W = len(windows)
        self.N = n_samples
        self.H = n_haps
        self.C = num_cycles
        self.bits = bits
        self.keep = []
        self.M = [int(w["tran"].shape[0]) for w in windows]
        wins = (GlpWindow * self.W)()
        self.parents_d, self.like_d = [], []
        for i, w in enumerate(windows):
            M = self.M[i]
            wn = (M + 63) // 64
            assert w["panel"].shape[1] == wn and w["panel"].shape[0] == n_haps
            per = 32 // bits
            assert w["packed"].shape == ((M + per - 1) // per, n_samples)
            panel = _dev(torch, w["panel"], self.device)
            packed = _dev(torch, w["packed"], self.device)
            tran = _dev(torch, np.asarray(w["tran"], np.float32), self.device)
            par = torch.zeros((n_samples, 4), dtype=torch.int32, device=self.device)
            if init_parents is not None:
                par.copy_(_dev(torch, np.asarray(init_parents[i], np.uint32), self.device))
            like = torch.zeros(n_samples, dtype=torch.float64, device=self.device)
            self.keep += [panel, packed, tran]
            self.parents_d.append(par)
            self.like_d.append(like)
            wins[i] = GlpWindow(M, wn, panel.data_ptr(), packed.data_ptr(), tran.data_ptr(),
                                par.data_ptr(), like.data_ptr())
        self.wins = wins
        ncode = 1 << bits
        cb = np.ascontiguousarray(codebook, np.float32).reshape(ncode, 2)
        pcm = np.ascontiguousarray(pc, np.float32).reshape(16)
        nbytes = lib().glp_emit_lut_bytes(bits, precision)
        self.lut = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
        check(lib().glp_build_emit_lut(ptr(cb, c_f32p), ncode, ptr(pcm, c_f32p), precision,
                                       self.lut.data_ptr()), "glp_build_emit_lut")
        d = GlpMhDesc()
        d.num_windows = self.W
        d.windows = ctypes.cast(wins, ctypes.POINTER(GlpWindow))
        d.num_samples = n_samples
        d.num_haps = n_haps
        d.bits_per_code = bits
        d.emit_lut = self.lut.data_ptr()
        d.num_cycles = num_cycles
        d.k0 = k0
        d.pen = 1.0
        d.seed = 0
        d.gen = 0
        d.prop_mode = prop_mode
        if klist is not None:
            self.klist_d = _dev(torch, np.asarray(klist, np.uint32), self.device)
            d.klist = self.klist_d.data_ptr()
            d.K = int(np.asarray(klist).shape[-1])
        if extra is not None:
            self.extra_d = _dev(torch, np.asarray(extra, np.uint32), self.device)
            d.extra = self.extra_d.data_ptr()
        d.precision = precision
        d.kickstart_from_ref = int(kickstart)
        self.trace_like = self.trace_acc = None
        if trace:
            self.trace_like = torch.full((self.W, n_samples, num_cycles + 1), float("nan"),
                                         dtype=torch.float64, device=self.device)
            self.trace_acc = torch.full((self.W, n_samples, max(num_cycles, 1)), 255,
                                        dtype=torch.uint8, device=self.device)
            d.trace_like = self.trace_like.data_ptr()
            d.trace_acc = self.trace_acc.data_ptr()
        self.desc = d
        self.plan = None
        self.extra_plans = []

    # -- execution ----------------------------------------------------------
    def stream_handle(self):
        return ctypes.c_void_p(self.torch.cuda.current_stream(self.device).cuda_stream)

    def make_plan(self, lanes=0):
        if self.plan is not None:
            check(lib().glp_plan_destroy(self.plan), "glp_plan_destroy")
        p = ctypes.c_void_p()
        check(lib().glp_plan_create(ctypes.byref(self.desc), lanes, ctypes.byref(p)), "glp_plan_create")
        self.plan = p
        return self.lanes()

    def autotune(self, gen=0, pen=1.0, seed=0):
        """Runs the library's autotuner; returns lanes per chain per window."""
        if self.plan is None:
            self.make_plan()
        out = np.zeros(self.W, np.uint32)
        check(lib().glp_plan_autotune(self.plan, gen, pen, seed, self.stream_handle(),
                                      ptr(out, c_u32p)), "glp_plan_autotune")
        return out

    def set_lanes(self, lanes):
        q = np.ascontiguousarray(np.broadcast_to(np.asarray(lanes, np.uint32), (self.W,)))
        check(lib().glp_plan_set_lanes(self.plan, ptr(q, c_u32p)), "glp_plan_set_lanes")

    def set_waves(self, waves):
        """Wavefronts per workgroup of the batched kernel (4 or 8)."""
        check(lib().glp_plan_set_waves(self.plan, waves), "glp_plan_set_waves")

    def waves(self):
        w = np.zeros(1, np.uint32)
        check(lib().glp_plan_waves(self.plan, ptr(w, c_u32p)), "glp_plan_waves")
        return int(w[0])

    def set_split(self, split):
        """Uneven chain slots at 8 wavefronts per workgroup: wavefronts 4-7
        hold one chain slot fewer than wavefronts 0-3 (results unchanged)."""
        check(lib().glp_plan_set_split(self.plan, 1 if split else 0), "glp_plan_set_split")

    def split(self):
        v = ctypes.c_int()
        check(lib().glp_plan_split(self.plan, ctypes.byref(v)), "glp_plan_split")
        return bool(v.value)

    def lanes(self):
        q, g = ctypes.c_uint32(), ctypes.c_uint32()
        check(lib().glp_plan_lanes(self.plan, ctypes.byref(q), ctypes.byref(g)), "glp_plan_lanes")
        return q.value, g.value

    def plan_with(self, prop_mode, lanes=0):
        """A second plan over the same device buffers with another proposal
        mode (e.g. uniform candidates before K-lists exist); released by close()."""
        d = GlpMhDesc.from_buffer_copy(self.desc)
        d.prop_mode = prop_mode
        p = ctypes.c_void_p()
        check(lib().glp_plan_create(ctypes.byref(d), lanes, ctypes.byref(p)), "glp_plan_create")
        self.extra_plans.append(p)
        return p

    def set_klist(self, klist):
        """Overwrite the K-lists in place ([N][K] for one window, or
        [W][N][K]; a device tensor or a host array of the shape given at
        construction)."""
        src = klist if hasattr(klist, "data_ptr") else _dev(self.torch, np.asarray(klist, np.uint32),
                                                           self.device)
        self.klist_d.copy_(src.reshape(self.klist_d.shape))

    def run_batched(self, gen=0, pen=1.0, seed=0, plan=None):
        if self.plan is None:
            self.make_plan()
        check(lib().glp_plan_run(plan or self.plan, gen, pen, seed, self.stream_handle()),
              "glp_plan_run")

    def run_plain(self, gen=0, pen=1.0, seed=0):
        if self.plan is None:
            self.make_plan()
        check(lib().glp_plan_run_plain(self.plan, gen, pen, seed, self.stream_handle()),
              "glp_plan_run_plain")

    def run_batched_graph(self, gen=0, pen=1.0, seed=0):
        """run_batched through the plan's HIP graph."""
        if self.plan is None:
            self.make_plan()
        check(lib().glp_plan_graph_run(self.plan, gen, pen, seed, self.stream_handle()),
              "glp_plan_graph_run")

    def run_plain_graph(self, gen=0, pen=1.0, seed=0):
        """run_plain through the plan's HIP graph."""
        if self.plan is None:
            self.make_plan()
        check(lib().glp_plan_graph_run_plain(self.plan, gen, pen, seed, self.stream_handle()),
              "glp_plan_graph_run_plain")

    def run_noalloc(self, gen=0, pen=1.0, seed=0, lanes=0, stats=False):
        """One batched generation through glp_mh_run_noalloc (no plan, no
        device allocation); with stats=True waits and returns (passes
        evaluated, useful passes, round-bound flag)."""
        self.desc.gen, self.desc.pen, self.desc.seed = gen, pen, seed
        out = np.zeros(3, np.uint64)
        check(lib().glp_mh_run_noalloc(ctypes.byref(self.desc), lanes, self.stream_handle(),
                                       ptr(out, c_u64p) if stats else None),
              "glp_mh_run_noalloc")
        return tuple(int(x) for x in out) if stats else None

    def run_plain_desc(self, gen=0, pen=1.0, seed=0):
        """The plain path through the stateless glp_plain_init / glp_plain_step."""
        self.desc.gen, self.desc.pen, self.desc.seed = gen, pen, seed
        s = self.stream_handle()
        check(lib().glp_plain_init(ctypes.byref(self.desc), s), "glp_plain_init")
        for c in range(self.C):
            check(lib().glp_plain_step(ctypes.byref(self.desc), c, s), "glp_plain_step")

    def stats(self):
        e, u = ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().glp_plan_stats(self.plan, ctypes.byref(e), ctypes.byref(u)), "glp_plan_stats")
        return e.value, u.value

    # -- results ------------------------------------------------------------
    def parents(self):
        self.torch.cuda.synchronize(self.device)
        return np.stack([p.cpu().numpy().view(np.uint32) for p in self.parents_d])

    def like(self):
        self.torch.cuda.synchronize(self.device)
        return np.stack([l.cpu().numpy() for l in self.like_d])

    def trace(self):
        self.torch.cuda.synchronize(self.device)
        return self.trace_like.cpu().numpy(), self.trace_acc.cpu().numpy()

    def set_panel(self, w, panel):
        """Replace window w's haplotype panel (e.g. after a generation)."""
        t = self.keep[3 * w]
        t.copy_(_dev(self.torch, panel, self.device))

    def close(self):
        if self.plan is not None:
            lib().glp_plan_destroy(self.plan)
            self.plan = None
        for p in getattr(self, "extra_plans", []):
            lib().glp_plan_destroy(p)
        self.extra_plans = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Resampler:
    """Impute::hmm_work on the device for the windows of a Job: after an MH
    generation (Job.run_batched / run_plain) it samples every sample's two
    new haplotypes from the chosen parents.

    het_aa: per window, the input GL pairs (Het, AA) as a host array
    [N][M][2] (Insti's prob layout); they are stored site-major on the device.
    pc: the 4 x 4 mutation matrix.
    emit_table: build each window's emission table once here (glp_work_emit,
    Insti::initialize's emit arrays) for every run to read; False leaves it
    to glp_hmm_work, which then rebuilds it in its workspace on every call.
    """

    def __init__(self, job, het_aa, pc, emit_table=True):
        torch = job.torch
        self.job = job
        self.torch = torch
        assert len(het_aa) == job.W
        self.q_d, self.hnew_d, self.emit_d = [], [], []
        pc16 = np.ascontiguousarray(np.asarray(pc, np.float32).reshape(16))
        wins = (GlpWorkWindow * job.W)()
        for w in range(job.W):
            M = job.M[w]
            wn = (M + 63) // 64
            q = np.asarray(het_aa[w], np.float32)
            assert q.shape == (job.N, M, 2)
            qd = _dev(torch, np.ascontiguousarray(q.transpose(1, 0, 2)), job.device)
            hn = torch.zeros((2 * job.N, wn), dtype=torch.int64, device=job.device)
            self.q_d.append(qd)
            self.hnew_d.append(hn)
            ep = None
            if emit_table:
                ed = torch.empty((M, job.N, 4), dtype=torch.float32, device=job.device)
                check(lib().glp_work_emit(qd.data_ptr(), M, job.N,
                                          pc16.ctypes.data_as(c_f32p), ed.data_ptr(),
                                          job.stream_handle()), "glp_work_emit")
                self.emit_d.append(ed)
                ep = ed.data_ptr()
            panel, tran = job.keep[3 * w], job.keep[3 * w + 2]
            wins[w] = GlpWorkWindow(M, wn, panel.data_ptr(), job.parents_d[w].data_ptr(),
                                    tran.data_ptr(), qd.data_ptr(), ep, hn.data_ptr())
        self.wins = wins
        d = GlpWorkDesc()
        d.num_windows = job.W
        d.windows = ctypes.cast(wins, ctypes.POINTER(GlpWorkWindow))
        d.num_samples = job.N
        d.num_haps = job.H
        for k, v in enumerate(np.asarray(pc, np.float32).reshape(16)):
            d.pc16[k] = float(v)
        nbytes = lib().glp_work_workspace_bytes(ctypes.byref(d))
        self.ws = torch.empty(max(nbytes, 16), dtype=torch.uint8, device=job.device)
        d.workspace = self.ws.data_ptr()
        self.desc = d

    def run(self, pen=1.0, seed=0, gen=0):
        self.desc.pen, self.desc.seed, self.desc.gen = pen, seed, gen
        check(lib().glp_hmm_work(ctypes.byref(self.desc), self.job.stream_handle()), "glp_hmm_work")

    def hnew(self):
        self.torch.cuda.synchronize(self.job.device)
        return np.stack([h.cpu().numpy().view(np.uint64) for h in self.hnew_d])

    def adopt(self):
        """Copies each window's new sample haplotypes into rows [0, 2N) of its
        panel (Insti's haps/hnew swap, insti.cpp:1455-1458); rows past 2N
        (reference haplotypes) are kept."""
        for w in range(self.job.W):
            self.job.keep[3 * w][: 2 * self.job.N].copy_(self.hnew_d[w])


def knn_lists(panel, num_sites, num_samples, K, freq_lb=0.0, freq_ub=1.0, using_maf=False,
              device="cuda:0"):
    """kNN candidate lists of samples 0 .. num_samples-1 over `panel` (a host
    [H][wn] uint64 array or a device tensor), built on the GPU; returns the
    device tensor [num_samples][glp_knn_k_out(K, H)] (int32 view of uint32)."""
    torch = _torch()
    dev = torch.device(device)
    pd = panel if hasattr(panel, "data_ptr") else _dev(torch, panel, dev)
    H = int(pd.shape[0])
    kout = lib().glp_knn_k_out(K, H)
    out = torch.zeros((num_samples, kout), dtype=torch.int32, device=dev)
    d = GlpKnnDesc(pd.data_ptr(), H, num_sites, num_samples, K, freq_lb, freq_ub,
                   int(using_maf), out.data_ptr(), None)
    nbytes = lib().glp_knn_workspace_bytes(ctypes.byref(d))
    ws = torch.empty(max(nbytes, 16), dtype=torch.uint8, device=dev)
    d.workspace = ws.data_ptr()
    s = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    check(lib().glp_knn_build(ctypes.byref(d), s), "glp_knn_build")
    torch.cuda.synchronize(dev)
    del ws
    return out
