"""Lifetime of contexts and their dependents (no GPU: a recording stand-in for
the library).  A batch or surrogate keeps raw pointers into its context, so the
library must see every dependent destroyed before the context; a dependent
closed (or collected) after its context must not reach the library again.
Closing a context with a live surrogate in the wrong order once ended a test
process with SIGSEGV inside dgv_sur_destroy."""
import ctypes as C
import gc

import numpy as np
import pytest

import dgv0d3v_direct_collision_amd.collision as col
import dgv0d3v_direct_collision_amd.surrogate as sur


class FakeLib:
    """Records destroy calls; hands out distinct handles."""

    def __init__(self):
        self.calls = []
        self.next = 100

    def _handle(self, out):
        self.next += 1
        out._obj.value = self.next
        return self.next

    def dgv_ctx_create(self, desc, device, out):
        self.calls.append(("ctx_create", self._handle(out)))
        return 0

    def dgv_ctx_num_nodes(self, ctx, n):
        n._obj.value = 27
        return 0

    def dgv_ctx_canonical_cell(self, ctx, cc):
        cc._obj.value = 0
        return 0

    def dgv_ctx_get_nodes(self, ctx, u, v, w, gw):
        return 0

    def dgv_ctx_destroy(self, ctx):
        self.calls.append(("ctx_destroy", ctx.value))

    def dgv_batch_create(self, arr, ngrids, goi, n, prec, out):
        self.calls.append(("batch_create", self._handle(out)))
        return 0

    def dgv_batch_destroy(self, b):
        self.calls.append(("batch_destroy", b.value))

    def dgv_sur_create(self, ctx, widths, nl, batch, seed, out):
        self.calls.append(("sur_create", self._handle(out)))
        return 0

    def dgv_sur_destroy(self, s):
        self.calls.append(("sur_destroy", s.value))


@pytest.fixture
def fake(monkeypatch):
    lib = FakeLib()
    monkeypatch.setattr(col, "load", lambda: lib)
    monkeypatch.setattr(sur, "load", lambda: lib)
    return lib


def _kinds(calls):
    return [k for k, _ in calls]


def test_dependents_destroyed_before_their_context(fake):
    g = col.DGV(1, 3, (-1.0, 1.0))
    b = col.DGVBatch([g], np.zeros(2, np.int32))
    s = sur.Surrogate(g, [27, 8, 27], batch=4)
    g.close()
    kinds = _kinds(fake.calls)
    assert kinds.index("ctx_destroy") > kinds.index("batch_destroy")
    assert kinds.index("ctx_destroy") > kinds.index("sur_destroy")
    n = len(fake.calls)
    b.close()            # already closed through the context
    s.close()
    g.close()            # twice is harmless
    assert len(fake.calls) == n


def test_collected_context_closes_its_dependents_first(fake):
    g = col.DGV(1, 3, (-1.0, 1.0))
    b = col.DGVBatch([g], np.zeros(2, np.int32))
    s = sur.Surrogate(g, [27, 8, 27], batch=4)
    del g                # the dependents still reference the context: nothing is freed yet
    gc.collect()
    assert "ctx_destroy" not in _kinds(fake.calls)
    del b, s
    gc.collect()
    kinds = _kinds(fake.calls)
    # every destroy happened once, and the context's last
    assert sorted(kinds) == sorted(["ctx_create", "batch_create", "sur_create",
                                    "batch_destroy", "sur_destroy", "ctx_destroy"])
    assert kinds[-1] == "ctx_destroy" or kinds.index("ctx_destroy") > max(
        kinds.index("batch_destroy"), kinds.index("sur_destroy"))


def test_dependent_closed_alone_leaves_the_context(fake):
    g = col.DGV(1, 3, (-1.0, 1.0))
    b = col.DGVBatch([g], np.zeros(2, np.int32))
    b.close()
    assert _kinds(fake.calls)[-1] == "batch_destroy"
    g.close()
    assert _kinds(fake.calls).count("batch_destroy") == 1
    assert _kinds(fake.calls)[-1] == "ctx_destroy"
