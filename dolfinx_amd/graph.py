"""Recording a sequence of library calls into one HIP graph.

Inside ``with g.capture():`` every kernel the library launches from this
thread is recorded instead of run; ``g.launch()`` then replays the whole
sequence as one graph launch on the library stream. It is meant for fixed
sequences run many times, such as many meshes assembled and solved at once:

    with g.capture():
        for A in mats: A.zero()
        fem.assemble_matrix_batched(forms, bcs, mats)
        ...vectors, BCs...
        batch.restart()
        batch.step(200)
    for coefficients in stream:
        ...set_coefficients on the forms...
        g.launch()

Replays read the device buffers the recorded calls used (coefficients,
matrices, vectors), which must stay alive at the same addresses; kernel
arguments, such as the constants of a form, keep their recorded values.
Calls that need the host part-way (allocations, plan or table uploads,
waits) raise DfxError with code DFX_ERR_CAPTURE inside a capture: run the
sequence once before recording it.
"""

from __future__ import annotations

import ctypes as C
from contextlib import contextmanager

from . import _lib


class Graph:
    def __init__(self):
        self._h = C.c_void_p()

    @contextmanager
    def capture(self):
        if self._h:
            raise RuntimeError("graph already recorded")
        _lib.call("dfx_graph_begin")
        try:
            yield self
        except BaseException:
            # close the capture (discarding it) and report the body's error
            h = C.c_void_p()
            _lib.lib().dfx_graph_end(C.byref(h))
            if h:
                _lib.lib().dfx_graph_destroy(h)
            raise
        _lib.call("dfx_graph_end", C.byref(self._h))

    def launch(self) -> None:
        """Replay the recorded sequence (asynchronous on the library stream)."""
        if not self._h:
            raise RuntimeError("nothing recorded")
        _lib.call("dfx_graph_launch", self._h)

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            try:
                _lib.lib().dfx_graph_destroy(h)
            except Exception:
                pass
