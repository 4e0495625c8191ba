"""wenofd_delta_t_noalloc (the allocation-free calculateDeltaT, partial
minima in a static device array) against wenofd_calculate_delta_t on the same
states: bit for bit, in both FP modes, on two calls in a row, the second with
a smaller grid; a NaN state raises the flag; a grid below the minimum is
refused. Device memory for the caller's side comes from torch (the caller owns
its buffers; the entry point allocates none)."""
import numpy as np
import pytest
import torch

import wenofd as W

pytestmark = pytest.mark.gpu


def _device_state(U, nx, ny, pitch):
    """Interior state [nx][ny][4] -> the library's layout [4][ny+8][pitch] on the GPU."""
    host = np.zeros((4, ny + 8, pitch))
    host[:, 4:ny + 4, 4:nx + 4] = np.transpose(U, (2, 1, 0))
    return torch.from_numpy(host).to("cuda")


def _noalloc_dt(s, U, fp, pitch):
    inf = s.info()
    dev = _device_state(U, inf.nx, inf.ny, pitch)
    out = torch.full((2,), -7.0, dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    W.delta_t_noalloc(dev.data_ptr(), inf.nx, inf.ny, pitch, inf.gamma, inf.dx, inf.dy, inf.cfl,
                      inf.t_end - inf.now, fp, out.data_ptr(), stream)
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("fp", [0, 1])
def test_noalloc_delta_t_equals_calculate_delta_t(fp):
    cases = [(dict(TESTCASE=5, QUAD_SEED=4, ELEMNUMX=300, ELEMNUMY=200), 3),
             (dict(TESTCASE=4, ELEMNUMX=100, ELEMNUMY=60), 2)]
    for keys, steps in cases:  # two calls in a row, the second on a smaller grid
        print(f"[gpu] delta_t_noalloc FP_CONTRACT={fp} {keys}", flush=True)
        with W.Solver(SCHEME=1, RKMETH=2, CFL=0.6, FP_CONTRACT=fp, **keys) as s:
            s.advance(steps)
            U = s.get_state()
            inf = s.info()
            got = _noalloc_dt(s, U, fp, inf.nx + 8 + 3)
            s.set_state(U)  # calculateDeltaT from the state itself, as the entry point computes it
            want = s.calculate_delta_t()
        assert got[1] == 0.0
        assert got[0] == want, (got[0], want)


def test_noalloc_delta_t_clips_and_flags_nan():
    with W.Solver(TESTCASE=2, SCHEME=1, RKMETH=2, CFL=0.6, ELEMNUMX=24, ELEMNUMY=24, TEND=1e-6) as s:
        U = s.get_state()
        got = _noalloc_dt(s, U, 0, 32)
        assert got[0] == 1e-6 and got[1] == 0.0  # clipped to t_left
        U[5, 7, 3] = -1.0  # negative energy: NaN sound speed
        got = _noalloc_dt(s, U, 0, 32)
        assert got[1] == 1.0
    with pytest.raises(W.WenofdError) as e:
        W.delta_t_noalloc(1, 4, 10, 16, 1.4, 0.1, 0.1, 0.5, 1.0, 0, 1)
    assert e.value.code == W.E_ARG
