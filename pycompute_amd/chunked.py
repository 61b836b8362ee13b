"""Chunked LM execution with convergence compaction (kernel pca_lm_chunk).

The iterations are split into chunks; after each chunk every wavefront moves
its still-active voxels into a packed queue with one atomic, so the next chunk
runs full wavefronts of unconverged voxels only. Per-voxel results are
identical to the single-launch kernel (tested on the GPU). The sequence of
chunk launches has a fixed shape, so it can be captured once in a hipGraph and
replayed.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from .lm import ptr


def chunk_schedule(iters: int, k: int, max_chunks: int = 0) -> list:
    """Chunks of k iterations covering `iters` (the last one shorter). With
    max_chunks > 0 the last of at most max_chunks chunks takes all remaining
    iterations: the voxels still active after the first few chunks mostly run
    to the iteration cap, so compacting them again only moves them."""
    if iters <= 0:
        return []
    k = max(1, int(k))
    out = [k] * (iters // k)
    if iters % k:
        out.append(iters % k)
    m = int(max_chunks)
    if m > 0 and len(out) > m:
        out = out[:m - 1] + [sum(out[m - 1:])]
    return out


class Workset:
    """Two packed queues (ping-pong) for one model's voxels. A queue entry is
    the state that changes (parameters, lambda, f, first f, step type,
    iteration count), the voxel's index and, in the carried layout, its data,
    bounds and per-voxel consts (ncm values per voxel), so that every later
    chunk reads them coalesced. carried=False leaves those out: the kernel then
    reads them from the caller's arrays through the index (less memory, more
    scattered reads)."""

    def __init__(self, N, n, dtype, device, M=0, ncm_per_voxel=0, carried=True):
        self.cap = max(int(N), 1)
        kw = dict(dtype=dtype, device=device)
        self.tensors = []
        self.structs = []
        for _ in range(2):
            t = {"y": torch.empty((M, self.cap), **kw) if carried else None,
                 "cv": (torch.empty((ncm_per_voxel, self.cap), **kw)
                        if carried and ncm_per_voxel else None),
                 "lo": torch.empty((n, self.cap), **kw) if carried else None,
                 "hi": torch.empty((n, self.cap), **kw) if carried else None,
                 "p": torch.empty((n, self.cap), **kw), "lam": torch.empty(self.cap, **kw),
                 "f": torch.empty(self.cap, **kw), "f_first": torch.empty(self.cap, **kw),
                 "idx": torch.empty(self.cap, dtype=torch.int32, device=device),
                 "cnt": torch.empty(self.cap, dtype=torch.int32, device=device),
                 "st": torch.empty(self.cap, dtype=torch.int8, device=device)}
            self.tensors.append(t)
            self.structs.append(_lib.WS(**{k: ptr(v) for k, v in t.items()}, cap=self.cap))


class ChunkedRun:
    """One or several models' fits as one chunked launch sequence.

    jobs: list of (program, LMArgs, uniform-consts tensor or None, Workset).
    All programs must be the same module (one pca_lm_chunk kernel).
    """

    def __init__(self, program, jobs, device, block: int, grid: int | None = None):
        self.program = program
        self.jobs = jobs
        self.block = block
        n_total = sum(int(a.N) for _, a, _, _ in jobs)
        # one wavefront task covers 64 * vpl queue slots
        slots = 64 * int(getattr(program.spec, "vpl", 1))
        waves = (n_total + slots - 1) // slots
        self.grid = grid or max(1, min((waves * 64 + block - 1) // block, 8192))
        arr = (_lib.ModelJob * len(jobs))()
        for i, (_, a, uc, ws) in enumerate(jobs):
            arr[i].a = a
            arr[i].uc = ptr(uc)
            arr[i].ws[0] = ws.structs[0]
            arr[i].ws[1] = ws.structs[1]
        self.job_array = arr
        self.device = device
        self._counts = None
        self._graph = None
        self._graph_key = None
        self._graph_counts = None

    def _counts_for(self, n_chunks):
        """Queue-length counters for a sequence of n_chunks chunks. A captured
        graph holds the address of the counters it was recorded with, so a
        longer launch after a capture gets a new buffer and the captured one
        stays alive (self._graph_counts) until the graph is destroyed."""
        need = (n_chunks + 1) * len(self.jobs)
        if self._counts is None or self._counts.numel() < need:
            self._counts = torch.zeros(need, dtype=torch.int32, device=self.device)
        return self._counts

    def set_iters(self, iters: int):
        for i in range(len(self.jobs)):
            self.job_array[i].a.iters = int(iters)

    def launch(self, schedule, stream=None, reset_counts=True, fresh=False):
        """Launch the chunk sequence. reset_counts=False: the counters are not
        reset first (possible for a single chunk, which reads no counter).
        fresh=True: the first chunk starts every voxel from the setup() state
        instead of reading it."""
        if not schedule:
            return
        counts = self._counts_for(len(schedule))
        s = torch.cuda.current_stream(self.device).cuda_stream if stream is None else stream
        k = (ctypes.c_int32 * len(schedule))(*schedule)
        L = _lib.lib()
        reset = self.program.kernel("pca_reset_counts") if reset_counts else None
        _lib.check(L.pca_lm_chunked_launch(self.program.kernel("pca_lm_chunk"), reset,
                                           self.job_array, len(self.jobs), counts.data_ptr(), k,
                                           len(schedule), self.grid, self.block, int(fresh), s),
                   "pca_lm_chunked_launch")

    # -- hipGraph -------------------------------------------------------------
    def capture(self, schedule, prologue=None, fresh=False):
        """Capture prologue() (e.g. the init kernel) and the chunk sequence into
        a graph on a side stream; returns the graph handle. fresh=True: every
        replay starts the fit from the setup() state (no reset kernel); a
        single chunk then needs no counter reset either."""
        self.destroy_graph()
        self._counts_for(len(schedule))
        self.program.kernel("pca_lm_chunk")
        self.program.kernel("pca_reset_counts")
        side = torch.cuda.Stream(self.device)
        side.wait_stream(torch.cuda.current_stream(self.device))
        L = _lib.lib()
        with torch.cuda.stream(side):
            h = side.cuda_stream
            _lib.check(L.pca_graph_begin(h), "pca_graph_begin")
            ok = False
            try:
                if prologue is not None:
                    prologue(h)
                self.launch(schedule, stream=h, reset_counts=not (fresh and len(schedule) == 1),
                            fresh=fresh)
                ok = True
            finally:
                g = ctypes.c_void_p()
                rc = L.pca_graph_end(h, ctypes.byref(g))
                if ok:
                    _lib.check(rc, "pca_graph_end")
        torch.cuda.current_stream(self.device).wait_stream(side)
        self._graph = g
        self._graph_key = tuple(schedule)
        self._graph_counts = self._counts      # the buffer the graph addresses
        return g

    def replay(self, stream=None):
        if self._graph is None:
            raise RuntimeError("no graph captured")
        s = torch.cuda.current_stream(self.device).cuda_stream if stream is None else stream
        _lib.check(_lib.lib().pca_graph_launch(self._graph, s), "pca_graph_launch")

    def destroy_graph(self):
        if self._graph is not None:
            _lib.lib().pca_graph_destroy(self._graph)
            self._graph = None
        self._graph_counts = None

    def __del__(self):
        try:
            self.destroy_graph()
        except Exception:
            pass
