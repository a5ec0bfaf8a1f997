"""Two-slot frame pipeline of the reference's interactive loop.

Renderer/OutputVisualizer.cpp:71-123 alternates frame slots 0 and 1: while
TCNNPass_p2 of frame f (the per-pair network) runs on the network's stream,
frame f + 1 is rasterised and encoded by TCNNPass_p1 on the render stream
into the other slot; TCNNSync then waits for p2 before the slot is reused.

``FramePipeline`` keeps that overlap with events instead of a full sync:

- the side stream updates frame f's environment map and runs p1 into slot
  f % 2, after p2 of frame f - 2 (the slot's and the map's last reader)
  has finished;
- the context stream waits for that p1 and runs p2 of frame f.

Frames alternate between two environment maps as well as two slots, so a
map update never races the p2 that reads the other map.  Each frame does
the same work as one sequential envmap update + relight.
"""
from __future__ import annotations

from typing import Optional, Sequence

from ._hip import DeviceBuffer, Event, Stream


class FramePipeline:
    """Submit frames; each runs map update + p1 on a side stream and p2 on
    the context stream, overlapping p2 of frame f with p1 of frame f + 1.

    ``envs`` are two faces-built environment maps of the context, one per
    slot (frame f uses ``envs[f % 2]``)."""

    def __init__(self, ctx, scene: int, envs: Sequence[int], precision: int,
                 tonemap: Optional[int] = None):
        from . import TONEMAP_FULL
        if len(envs) != 2:
            raise ValueError("FramePipeline needs one environment map per slot (two)")
        self.ctx = ctx
        self.scene = scene
        self.envs = list(envs)
        self.precision = precision
        self.tonemap = TONEMAP_FULL if tonemap is None else tonemap
        self.side = Stream()
        self.p1_done = [Event(), Event()]
        self.p2_done = [Event(), Event()]
        self.n = 0

    def submit(self, frame, faces: Optional[DeviceBuffer] = None) -> int:
        """Queue the next frame; returns its slot.  ``faces`` (device) are
        that frame's cubemap; None keeps the slot's map as it is."""
        s = self.n % 2
        side, net = self.side.handle, self.ctx.stream
        if self.n >= 2:
            self.p2_done[s].wait(side)
        if faces is not None:
            self.ctx.envmap_update_faces(self.envs[s], faces, side)
        self.ctx.TCNNPass_p1(self.scene, s, frame, self.precision, side)
        self.p1_done[s].record(side)
        self.p1_done[s].wait(net)
        self.ctx.TCNNPass_p2(self.scene, self.envs[s], s, frame, self.precision, self.tonemap,
                             net)
        self.p2_done[s].record(net)
        self.n += 1
        return s

    def drain(self) -> None:
        """Wait for every submitted frame."""
        self.side.synchronize()
        self.ctx.sync()
        self.n = 0
