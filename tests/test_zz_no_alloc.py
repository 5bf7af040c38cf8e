"""osb_step_noalloc -- allocation-free stepping of a caller-owned block --
against osb_step through the C ABI.  It runs the same kernels on buffers
carved from the library's static scratch array, so the padded states agree
bit for bit: two calls in a row, the second on a smaller block (the scratch
keeps the first call's contents) and over an odd number of stages (the result
is copied back from the scratch), for a shock-capturing 2-D block, a 1-D one
with CFL steps and a fused 3-D central block; a block beyond the scratch is
refused with OSB_E_UNSUPPORTED before anything is launched."""
import numpy as np
import pytest

from opensbli_amd import native as nat, problems as P

pytestmark = pytest.mark.gpu


def _announce(name):
    print(f"[gpu-case] {name}", flush=True)


def _sod_cfl(n, steps=40, cfl=0.4):
    pr = P.sod(n=n, steps=steps)
    pr.desc.cfl, pr.desc.dt = cfl, 0.0
    pr.name += "_cfl"
    return pr


def _reference(hip, pr, steps):
    """Padded initial state (halos filled by the context) and padded state after `steps` via osb_step."""
    with hip.context(pr.desc) as c:
        c.set_state(pr.q0)
        q0p = c.get_state(nat.LAYOUT_PADDED)
        c.step(steps)
        c.sync()
        return q0p, c.get_state(nat.LAYOUT_PADDED)


def _noalloc(hip, pr, q0p, steps, torch):
    q = torch.from_numpy(np.ascontiguousarray(q0p).reshape(-1).copy()).to("cuda")
    du = torch.zeros_like(q)
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()
    hip.step_noalloc(pr.desc, q.data_ptr(), du.data_ptr(), steps, stream.cuda_stream)
    stream.synchronize()
    return q.cpu().numpy().reshape(q0p.shape)


@pytest.mark.parametrize("pair", [
    (lambda: P.kelvin_helmholtz(n=64, steps=12), lambda: P.kelvin_helmholtz(n=48, steps=7)),
    (lambda: _sod_cfl(512), lambda: _sod_cfl(256, steps=21)),
    (lambda: P.tgv(n=24, steps=4), lambda: P.tgv(n=16, steps=3)),
], ids=["kh", "sod_cfl", "tgv_fused"])
def test_noalloc_equals_step_twice_then_smaller(hip, pair):
    torch = pytest.importorskip("torch")
    assert torch.cuda.is_available()
    for make in pair:
        pr = make()
        _announce(f"noalloc {pr.name} n={list(pr.desc.n)} x {pr.steps}")
        q0p, want = _reference(hip, pr, pr.steps)
        got = _noalloc(hip, pr, q0p, pr.steps, torch)
        assert np.array_equal(got, want), float(np.abs(got - want).max())


def test_noalloc_refuses_a_block_beyond_the_scratch(hip):
    torch = pytest.importorskip("torch")
    pr = P.tgv(n=64, steps=1)                     # 5 x 68^3 doubles of ping-pong state alone exceed 8 MiB
    _announce("noalloc oversized block")
    q = torch.zeros(8, dtype=torch.float64, device="cuda")
    with pytest.raises(nat.OSBError) as e:
        hip.step_noalloc(pr.desc, q.data_ptr(), q.data_ptr(), 1, 0)
    assert e.value.code == nat.E_UNSUPPORTED
    assert "scratch" in str(e.value)
    # the library keeps working on the same device afterwards
    small = P.kelvin_helmholtz(n=32, steps=3)
    q0p, want = _reference(hip, small, 3)
    assert np.array_equal(_noalloc(hip, small, q0p, 3, torch), want)
