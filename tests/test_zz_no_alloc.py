"""lepic_efield_noalloc (the E field with its wall-pass scratch in a static
device array) against lepic_efield, both through the C ABI: bit-identical E
on two calls in a row, the second on a smaller grid (the static array keeps
the first call's values), and the size limit rejected before any launch."""
import ctypes as C

import numpy as np
import pytest
import torch

from essaitemporaire_lepic_module_amd import _native
from essaitemporaire_lepic_module_amd import config as cfg
from essaitemporaire_lepic_module_amd.device import context, ptr
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _both(dev, n, h, bc, phi):
    ctx = context(dev)
    g = _native.grid(n, h)
    phi_d = torch.as_tensor(phi, device=dev)
    bc_d = torch.as_tensor(bc, device=dev)
    shape = cfg.node_shape(n) + (4,)
    # the fourth component is carried through untouched: give it values
    w = torch.as_tensor(np.random.default_rng(1).standard_normal(shape[:-1]), device=dev)
    E_ref = torch.zeros(shape, dtype=torch.float64, device=dev)
    E_na = torch.zeros(shape, dtype=torch.float64, device=dev)
    E_ref[..., 3] = w
    E_na[..., 3] = w
    ctx.check(ctx.lib.lepic_efield(ctx.h, C.byref(g), ptr(phi_d), ptr(bc_d), ptr(E_ref)))
    st = ctx.lib.lepic_efield_noalloc(ctx.h, C.byref(g), ptr(phi_d), ptr(bc_d), ptr(E_na))
    assert st == 0, ctx.lib.lepic_last_error(ctx.h)
    torch.cuda.synchronize()
    return E_ref.cpu().numpy(), E_na.cpu().numpy()


def test_noalloc_efield_equals_efield(gpu, oracle, rng):
    # first call: the 128^3 workload's grid (131^3 nodes) with electrode
    # plates one plasma node apart, so the wall pass follows chains
    n, h = H.box((128, 128, 128))
    bc = H.thin_gap_bcnd(n)
    print(f"[no_alloc] efield {n} thin-gap plates", flush=True)
    E_ref, E_na = _both(gpu, n, h, bc, rng.standard_normal(cfg.node_shape(n)))
    assert np.array_equal(E_ref, E_na)
    assert np.isfinite(E_na).all()
    # second call, smaller: every node the pass sets is written again, none of
    # the first call's values leak in; both equal the oracle
    n, h = H.box((12, 10, 8))
    bc = H.thin_gap_bcnd(n)
    phi = rng.standard_normal(cfg.node_shape(n))
    print(f"[no_alloc] efield {n} thin-gap plates", flush=True)
    E_ref, E_na = _both(gpu, n, h, bc, phi)
    assert np.array_equal(E_ref, E_na)
    Eo = oracle.efield(n, h, phi, bc)
    assert np.array_equal(E_na[..., :3], Eo[..., :3])


def test_noalloc_efield_rejects_grids_above_the_limit(gpu):
    n, h = H.box((131, 131, 131))          # 134^3 nodes > LEPIC_EFIELD_NOALLOC_MAX_NODES
    assert (n[0] + 3) ** 3 > 2359296
    ctx = context(gpu)
    g = _native.grid(n, h)
    shape = cfg.node_shape(n)
    phi = torch.zeros(shape, dtype=torch.float64, device=gpu)
    bc = torch.as_tensor(cfg.grounded_box_bcnd(n), device=gpu)
    E = torch.zeros(shape + (4,), dtype=torch.float64, device=gpu)
    st = ctx.lib.lepic_efield_noalloc(ctx.h, C.byref(g), ptr(phi), ptr(bc), ptr(E))
    assert _native.ERRORS.get(st) == "ERR_ARG"
    assert b"exceed the static scratch" in ctx.lib.lepic_last_error(ctx.h)
