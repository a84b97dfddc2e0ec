"""The C-ABI library loads and exports every symbol include/tlbm_hip.h declares (no GPU work)."""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import pytest

from tensorlbm_amd import _native as nat

HEADER = Path(__file__).resolve().parent.parent / "include" / "tlbm_hip.h"


def declared() -> list[str]:
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tlbm_[a-z_0-9]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared() == nat.exported_symbols()


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(nat.library_path()))
    for name in declared():
        assert hasattr(lib, name), name


def test_version_and_status_strings():
    assert nat.lib().tlbm_version() >= 1
    assert nat.lib().tlbm_status_string(nat.ERR_ALIAS) == b"input and output alias"


def test_argument_validation_happens_before_any_launch():
    """Validation paths return error codes without touching a device."""
    lib = nat.lib()
    coll = nat.make_collision(nat.COLL_BGK, 0.7)
    p, p2 = ctypes.c_void_p(0x1000), ctypes.c_void_p(0x2000)

    def step(q=19, store=0, a=p, b=p2, dims=(4, 4, 4), c=coll, si=None, so=None):
        return lib.tlbm_step(q, store, a, si, b, so, *dims, c, None, None, None, None)

    assert step(b=p) == nat.ERR_ALIAS
    assert step(a=None) == nat.ERR_INVALID
    assert step(q=20) == nat.ERR_INVALID
    assert step(q=9, dims=(2, 4, 4)) == nat.ERR_SHAPE
    assert step(store=9) == nat.ERR_DTYPE
    assert step(c=nat.make_collision(nat.COLL_BGK, 0.4)) == nat.ERR_INVALID
    big = 1 << 11
    assert step(dims=(big, big, big)) == nat.ERR_SHAPE
    assert step(si=nat.Strides(64, 16, 2)) == nat.ERR_SHAPE  # y stride below nx
    # a z-plane spanning more than 4 GiB cannot use 32-bit lane offsets
    assert step(dims=(1, 40000, 40000), si=nat.Strides(1 << 40, 1 << 40, 40000)) in (nat.ERR_SHAPE,)
    assert lib.tlbm_stream(19, 0, p, None, p, None, 4, 4, 4, None) == nat.ERR_ALIAS
    assert lib.tlbm_collide(19, 2, p, None, p, None, 4, 4, 4, coll, None) == nat.ERR_DTYPE
    assert b"float32 or float64" in lib.tlbm_last_error()


def test_check_maps_status_to_reference_exceptions():
    with pytest.raises(ValueError):
        nat.check(nat.ERR_SHAPE, "x")
    with pytest.raises(NotImplementedError):
        nat.check(nat.ERR_UNSUPPORTED, "x")
    with pytest.raises(RuntimeError):
        nat.check(nat.ERR_HIP, "x")


def test_store_codes():
    import torch

    assert nat.store_code(torch.float16, deviation=True) == nat.STORE_F16_DEV
    assert nat.store_code(torch.bfloat16) == nat.STORE_BF16
    with pytest.raises(ValueError, match="Unsupported dtype"):
        nat.store_code(torch.int32)
    with pytest.raises(ValueError):
        nat.store_code(torch.float32, deviation=True)


def test_cavity_layout_is_host_only_and_validates():
    lib = nat.lib()
    tab = (nat.Cavity * 3)()
    for i, e in enumerate((8, 33, 5)):
        tab[i].edge = e
        tab[i].tau = 0.6
    tot, tiles = ctypes.c_int64(), ctypes.c_int64()
    assert lib.tlbm_cavities_layout(tab, 3, ctypes.byref(tot), ctypes.byref(tiles)) == nat.OK
    planes = [(e * e + 63) // 64 * 64 for e in (8, 33, 5)]
    assert [tab[i].offset for i in range(3)] == [0, 8 * 19 * planes[0], 8 * 19 * planes[0] + 33 * 19 * planes[1]]
    assert tot.value == sum(e * 19 * p for e, p in zip((8, 33, 5), planes))
    assert [tab[i].tiles_before for i in range(3)] == [0, 8, 8 + 33 * 5]
    assert tiles.value == 8 + 33 * 5 + 5
    p, p2 = ctypes.c_void_p(0x1000), ctypes.c_void_p(0x100000)
    step = lib.tlbm_cavities_step
    assert step(0, p, p, tot.value, tab, 3, 0.1, 0, None) == nat.ERR_ALIAS
    assert step(0, p, p2, tot.value - 1, tab, 3, 0.1, 0, None) == nat.ERR_SHAPE
    assert step(2, p, p2, tot.value, tab, 3, 0.1, 0, None) == nat.ERR_DTYPE
    assert step(0, p, p2, tot.value, tab, 3, 0.1, -1, None) == nat.ERR_INVALID
    tab[1].tau = 0.5
    assert step(0, p, p2, tot.value, tab, 3, 0.1, 0, None) == nat.ERR_INVALID
    tab[1].tau = 0.6
    tab[2].tiles_before += 1
    assert step(0, p, p2, tot.value, tab, 3, 0.1, 0, None) == nat.ERR_INVALID
    bad = (nat.Cavity * 1)()
    bad[0].edge = 2
    assert lib.tlbm_cavities_layout(bad, 1, None, None) == nat.ERR_SHAPE
    h = ctypes.c_void_p()
    assert lib.tlbm_cavities_graph(0, p, p2, tot.value, tab, 3, 0.1, 0, 3, ctypes.byref(h)) == nat.ERR_INVALID
    assert lib.tlbm_graph_destroy(None) == nat.OK


def test_mlp_sizes_without_a_device():
    v = [ctypes.c_int64() for _ in range(4)]
    assert nat.lib().tlbm_mlp_sizes(*[ctypes.byref(x) for x in v]) == nat.OK
    assert [x.value for x in v] == [1345, 1359, 3520, 64]


def test_noalloc_training_step_rejects_oversized_batches_before_any_launch():
    lib = nat.lib()
    p = ctypes.c_void_p(0x1000)
    args = (p, p, 1, 1e-3, 0.9, 0.999, 1e-8, 1.0, None, None, None)
    assert lib.tlbm_mlp_train_step_noalloc(p, p, p, nat.MLP_NOALLOC_MAX_SAMPLES + 1, *args) == nat.ERR_SHAPE
    assert b"no-allocation limit" in lib.tlbm_last_error()
    assert lib.tlbm_mlp_train_step_noalloc(p, p, p, 0, *args) == nat.ERR_SHAPE
    assert lib.tlbm_mlp_train_step_noalloc(None, p, p, 64, *args) == nat.ERR_INVALID


def test_mlp2d_architecture_checks_without_a_device():
    from tensorlbm_amd import closure2d as c2

    lib = nat.lib()
    assert lib.tlbm_mlp2d_nparam(ctypes.byref(c2.ModelArch().native())) == 3 * 16 + 16 + 16 * 16 + 16 + 16 + 1
    assert lib.tlbm_mlp2d_nparam(ctypes.byref(c2.ModelArch(5, 64, 1, "gelu").native())) == 5 * 64 + 64 + 64 + 1
    assert lib.tlbm_mlp2d_nparam(ctypes.byref(c2.ModelArch(hidden_features=20).native())) == -1
    assert b"hidden_features" in lib.tlbm_last_error()
    assert lib.tlbm_mlp2d_nparam(ctypes.byref(c2.ModelArch(n_hidden_layers=5).native())) == -1
    with pytest.raises(ValueError, match="Unsupported activation"):
        c2.ModelArch(activation="swish").native()
    p = ctypes.c_void_p(0x1000)
    arch = c2.ModelArch(3, 64, 2).native()  # 64 x 2 > 120: training's LDS would exceed 64 KB
    assert lib.tlbm_mlp2d_grad(ctypes.byref(arch), p, p, p, 100, p, None) == nat.ERR_UNSUPPORTED
