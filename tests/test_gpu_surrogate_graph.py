"""The surrogate iteration as device work only and as a replayed graph:
the library's own network pass gives the predictions of the Python-driven
one bit for bit (fused width-256 kernel and the unfused width-512 path), the
device-only solve equals the Python-driven surrogate solve, and a captured
iteration with a 1 % Vs update replays, replay after replay with eager work
in between, to the eager solve of the same updated model."""
import numpy as np
import pytest

from jdsurfg_amd import workloads as W

pytestmark = pytest.mark.gpu
KERN = ("dcdb", "dudb")
NAMES = ("c", "u", "status") + KERN


def _setup(w, weights):
    import os
    from jdsurfg_amd import swd, surrogate as S
    path = os.path.join(os.path.dirname(S.DEFAULT_WEIGHTS), weights)
    mlp = S.MLP.load(path)
    plan = swd.Plan(w.layer_off, w.per_off, w.nmode, w.group, "csr")
    plan.set_tolerance(1e-10)
    return plan, mlp


@pytest.mark.parametrize("weights", ["surrogate_mlp.npz", "surrogate_mlp_512.npz"])
def test_device_only_iteration_matches_python_driven(weights):
    import torch
    from jdsurfg_amd import swd, surrogate as S
    w = W.uneven(ncol=300, seed=41)
    print(f"case device-only surrogate iteration, {weights}, uneven ncol=300", flush=True)
    plan, mlp = _setup(w, weights)
    inp = swd.upload(w)
    a = plan.alloc_outputs(KERN)
    S.surrogate_solve(plan, inp, a, mlp=mlp)
    torch.cuda.synchronize()
    pred_py = S.predict(mlp, plan._feats, e0=0, n=plan.nent).cpu().numpy()
    S.bind_network(plan, mlp)
    b = plan.alloc_outputs(KERN)
    S.surrogate_solve_net(plan, inp, b)
    torch.cuda.synchronize()
    assert np.array_equal(S.network_pred(plan), pred_py)
    for n in NAMES:
        assert torch.equal(a[n], b[n]), n


def test_captured_surrogate_iteration_replays_to_eager():
    import torch
    from jdsurfg_amd import swd, surrogate as S
    w = W.uneven(ncol=250, seed=42)
    print("case captured surrogate iteration, uneven ncol=250, 3 replays", flush=True)
    plan, mlp = _setup(w, "surrogate_mlp.npz")
    S.bind_network(plan, mlp)
    inp = swd.upload(w)
    out = plan.alloc_outputs(KERN)
    st = torch.cuda.Stream()
    with torch.cuda.stream(st):
        S.surrogate_solve_net(plan, inp, out, stream=st.cuda_stream)
        st.synchronize()
        g = S.capture_surrogate(plan, inp, out, st, perturb_amp=0.01, seed=5)
        # an eager plan on its own buffers gives each replay's reference
        eplan, _ = _setup(w, "surrogate_mlp.npz")
        S.bind_network(eplan, mlp)
        ref = eplan.alloc_outputs(KERN)
        prev = inp[2].clone()
        for it in range(3):
            g.launch(st)
            st.synchronize()
            vs = inp[2]
            ratio = (vs / torch.where(prev > 0, prev, torch.ones_like(prev)) - 1).abs().max()
            assert 0 < float(ratio) <= 0.0101
            prev = vs.clone()
            model = tuple(t.clone() for t in inp)
            S.surrogate_solve_net(eplan, model, ref, stream=st.cuda_stream)
            st.synchronize()
            for n in NAMES:
                assert torch.equal(out[n], ref[n]), (it, n)
            # eager work on the graph's plan between replays
            S.surrogate_solve_net(plan, model, plan.alloc_outputs(KERN), stream=st.cuda_stream)
            st.synchronize()
