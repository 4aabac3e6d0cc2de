"""Handle lifetimes of the Python binding, on the CPU (no GPU, no library calls).

A test process once ended with SIGSEGV after a failed GPU test: pytest -x
stopped with graphs still referenced from the failed test's frame, and the
objects were freed by the cyclic garbage collector at exit.  The context then
tracked its objects in a weak set.  The collector clears every weak reference
into a garbage cycle before it runs the finalizers of the cycle, so the set
was already empty when the context's finalizer asked whether objects were
left; the context destroyed its native handle, and the graphs' finalizers
then called vnm_graph_destroy on graphs whose context was freed.  The first
test shows that mechanism with plain Python objects; the second runs the
binding's own Context and Graph classes against a recording stand-in for the
library and checks that the context's handle goes last.
"""
import gc
import weakref

import vnm


class _Obj:
    pass


def test_weak_set_empties_before_finalizers_of_a_cycle():
    seen = []

    class Node:
        def __init__(self, registry):
            self.registry = registry
            registry.add(self)

        def __del__(self):
            seen.append(len(self.registry))

    reg = weakref.WeakSet()
    a, b = Node(reg), Node(reg)
    a.peer, b.peer = b, a  # only the cyclic collector frees these
    assert len(reg) == 2
    del a, b
    gc.collect()
    # both finalizers ran, and each saw the weak set empty already: a context
    # that counted its live objects this way would have freed its handle first
    assert seen == [0, 0]


class _RecordingLib:
    def __init__(self):
        self.calls = []

    def vnm_graph_destroy(self, h):
        self.calls.append(("graph", h))

    def vnm_ctx_destroy(self, h):
        self.calls.append(("ctx", h))


def test_context_handle_outlives_a_garbage_cycle_of_its_graphs(monkeypatch):
    rec = _RecordingLib()
    monkeypatch.setattr(vnm, "_lib", rec)
    ctx = object.__new__(vnm.Context)
    ctx.h, ctx._live, ctx._closing = "ctx-handle", 0, False
    graphs = []
    for i in range(2):
        g = object.__new__(vnm.Graph)
        g.h, g.ctx = f"graph-{i}", ctx
        ctx._adopt(g)
        graphs.append(g)
    graphs[0].peer, graphs[1].peer = graphs[1], graphs[0]
    ctx.close()
    assert rec.calls == [] and ctx.h == "ctx-handle"  # two graphs are alive
    del graphs, g
    gc.collect()
    assert sorted(rec.calls[:2]) == [("graph", "graph-0"), ("graph", "graph-1")]
    assert rec.calls[2:] == [("ctx", "ctx-handle")]
    assert ctx.h is None
