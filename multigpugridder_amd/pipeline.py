"""Device-resident projection pipeline over a fixed direction set.

One step is the plain drop-in workload with every input resident in HBM:
a check of the CAS volume and table (ops.check_volume, which lets the forward
gather skip the stencil taps whose weight is 0), forward projection of all
directions through to images
(``MultiGPUGridder_Matlab_Class.m:198-243``: gather, then CASImgsToImgs), then
back projection of those images into a zeroed CAS volume (``:245-270``:
ImgsToCASImgs, then the gather's reference-semantics transpose).  The
directions are processed in chunks spread over streams, so that the image
chain of one chunk runs beside the projector of the next; projector launches
are ordered among themselves (back projection read-modify-writes one volume).

Capturing the step into a hipGraph is switched off: a capture of it
(``torch.cuda.CUDAGraph`` over the pipeline's two streams, after an eager
warm-up) ended in a host segmentation fault inside the runtime's end of
capture, with no cause found in this code, so :meth:`capture` raises.
"""
from __future__ import annotations

import torch

from . import ops


class ProjectionPipeline:
    def __init__(self, cas_volume: torch.Tensor, axes: torch.Tensor, n: int, chunk: int = 2000,
                 num_streams: int = 2, ker: torch.Tensor | None = None):
        if cas_volume.dim() != 3 or axes.dim() != 2 or axes.shape[1] != 9:
            raise ValueError("cas_volume must be (V, V, V) and axes (P, 9)")
        self.dev = cas_volume.device
        self.cas_vol = cas_volume
        self.axes = axes
        self.n = int(n)
        self.m = 2 * self.n
        self.v = cas_volume.shape[-1]
        self.p = axes.shape[0]
        self.chunk = min(int(chunk), self.p)
        self.ker = ops.kb_table(self.dev) if ker is None else ker
        # default priority streams (the whole step runs on them)
        self.streams = [torch.cuda.Stream(device=self.dev) for _ in range(max(1, int(num_streams)))]
        self.cas_bufs = [torch.empty((self.chunk, self.m, self.m), device=self.dev) for _ in self.streams]
        self.imgs = torch.empty((self.p, self.n, self.n), device=self.dev)
        self.bp_vol = torch.zeros((self.v, self.v, self.v), device=self.dev)
        self.check = torch.empty(1, device=self.dev, dtype=torch.int32)
        self.graph = None

    def chunks(self):
        return [(s0, min(self.chunk, self.p - s0)) for s0 in range(0, self.p, self.chunk)]

    def step(self, timed=None):
        """Queues one step on the pipeline's streams (joined back into the
        current stream).  ``timed(name, fn, stream)`` may wrap each launch
        group ("fp", "fp_chain", "bp_chain", "bp"), e.g. with HIP events."""
        run = timed if timed is not None else (lambda name, fn, st: fn())
        main = torch.cuda.current_stream(self.dev)
        # the volume and table check behind the forward gather's shortcut,
        # once per step (the volume may have changed since the last one)
        ops.check_volume(self.cas_vol, self.ker, out=self.check)
        go = torch.cuda.Event()
        go.record(main)
        for st in self.streams:
            st.wait_event(go)
        ns = len(self.streams)
        prev = None
        m, n, v = self.m, self.n, self.v
        for k, (s0, c) in enumerate(self.chunks()):
            ax = self.axes[s0:s0 + c]
            st, buf = self.streams[k % ns], self.cas_bufs[k % ns]
            with torch.cuda.stream(st):
                if prev is not None:
                    st.wait_event(prev)  # projector launches do not overlap each other
                run("fp", lambda: ops.forward_project(self.cas_vol, ax, m, ker=self.ker, out=buf[:c],
                                                      check=self.check), st)
                prev = torch.cuda.Event()
                prev.record(st)
                run("fp_chain", lambda: ops.cas_images_to_images(buf[:c], n, out=self.imgs[s0:s0 + c],
                                                                 clobber=True), st)
        with torch.cuda.stream(self.streams[0]):
            self.bp_vol.zero_()
        for k, (s0, c) in enumerate(self.chunks()):
            ax = self.axes[s0:s0 + c]
            st, buf = self.streams[k % ns], self.cas_bufs[k % ns]
            with torch.cuda.stream(st):
                run("bp_chain", lambda: ops.images_to_cas_images(self.imgs[s0:s0 + c], m, out=buf[:c]), st)
                if prev is not None:
                    st.wait_event(prev)
                run("bp", lambda: ops.back_project(buf[:c], ax, v, ker=self.ker, out=self.bp_vol), st)
                prev = torch.cuda.Event()
                prev.record(st)
        for st in self.streams:
            done = torch.cuda.Event()
            done.record(st)
            main.wait_event(done)

    def capture(self, warmup: int = 1):
        """Switched off (see the module docstring): raises RuntimeError."""
        raise RuntimeError("hipGraph capture of the projection pipeline is switched off: a capture ended in a "
                           "host segmentation fault at the end of capture, cause not found")

    def replay(self):
        raise RuntimeError("hipGraph capture of the projection pipeline is switched off")
