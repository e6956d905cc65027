"""The learned warm start end to end on a small cylinder: steady states from
the solver, a coordinate MLP trained on them by the device's BF16 MFMA
kernels, and a held-out case started from the network's Maxwellian.  The
unmodified solver judges both runs by the same normaliser (the freestream
run's step-1 change); the warm start must reach the steady state in fewer
steps and to the same drag.  Also the shipped particle network on a held-out
particle case."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_trained_network_shortens_the_run_to_steady_state(tmp_path):
    from dgbgk import meshgen, warmstart
    elems, rorder, torder, tol = 1200, 8, 16, 1e-5
    train_set = [warmstart.Sample(m, k) for m in (0.5, 0.75, 1.0) for k in (0.1, 0.3)]
    held_out = warmstart.Sample(0.65, 0.18)
    mesh = warmstart.mesh_for(elems)
    wd = meshgen.wall_distance(mesh)
    print(f"warm start: mesh E={mesh.nelem}, {len(train_set)} training states", flush=True)
    X, Y = [], []
    for s in train_set:
        r = warmstart.run_to_steady(warmstart.cylinder_case(mesh, s, rorder, torder), tol,
                                    60000)
        assert r["converged"], (s, r["steps"])
        X.append(warmstart.features(mesh, wd, s))
        Y.append(warmstart.targets(r["fields"], s))
    from dgbgk.mlp import MLP
    m = MLP(fin=5, width=128, hidden=4, fout=4, batch=4096)
    m.init(seed=0)
    m.set_data(np.concatenate(X), np.concatenate(Y))
    losses = np.concatenate([m.train(500, lr=2e-3 if i < 4 else 5e-4, seed=i) for i in range(6)])
    print(f"training loss {losses[:20].mean():.3e} -> {losses[-50:].mean():.3e}", flush=True)
    assert losses[-50:].mean() < 0.1 * losses[:20].mean()

    case = warmstart.cylinder_case(mesh, held_out, rorder, torder)
    cold = warmstart.run_to_steady(case, tol, 60000, ref_tol=1e-8)
    # the run went on past tol to the reference tolerance; fields and drag
    # are those at tol, ref_drag the converged one
    assert cold["ref_steps"] is not None and cold["ref_steps"] > cold["steps"]
    assert abs(cold["ref_drag"] - cold["drag"]) <= 2e-3 * abs(cold["ref_drag"])
    pred = warmstart.untarget(m.forward(warmstart.features(mesh, wd, held_out)), held_out)
    warm = warmstart.run_to_steady(case, tol, 60000, seed=pred, norm=cold["first_change"])
    print(f"held out: cold {cold['steps']} steps, warm {warm['steps']} steps, "
          f"drag {cold['drag']:.6f} / {warm['drag']:.6f}", flush=True)
    assert cold["converged"] and warm["converged"]
    assert warm["steps"] <= 0.85 * cold["steps"]
    assert abs(warm["drag"] - cold["drag"]) <= 2e-3 * abs(cold["drag"])
    # the prediction is close to the steady density, not just to the freestream
    rho = cold["fields"]["rho"]
    assert np.linalg.norm(pred[0] - rho) < 0.5 * np.linalg.norm(1.0 - rho)
    m.close()


def test_committed_particle_network_shortens_the_run():
    """The shipped particle networks on the particle box (25,781 triangles,
    8 x 16 ordinates, Mach 0.3, Kn 1, a held-out case): FP64 from the
    network's state reaches a step change of 1e-6 of the freestream run's
    step 1 in at most half the steps (measured: 20,429 against 114,584)."""
    import os
    from dgbgk import _lib, meshgen, warmstart
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    params = os.path.join(root, "dg-bgk_amd", "dgbgk", "data", "warmstart_particle_8x16.npz")
    mesh = warmstart.mesh_for(23_000, "particle")
    smp = warmstart.Sample(0.3, 1.0)
    case = warmstart.cylinder_case(mesh, smp, 8, 16)
    print(f"particle E={mesh.nelem} V={case.nvel}: cold and warm FP64 to 1e-6", flush=True)
    m = warmstart.load_model(params)
    pred = warmstart.untarget(
        m.forward(warmstart.features(mesh, meshgen.wall_distance(mesh), smp, "particle")), smp)
    m.close()
    cold = warmstart.run_to_steady(case, 1e-6, 200_000, _lib.DGBGK_FP64)
    warm = warmstart.run_to_steady(case, 1e-6, 200_000, _lib.DGBGK_FP64, pred,
                                   cold["first_change"])
    print(f"cold {cold['steps']} steps, warm {warm['steps']} steps", flush=True)
    assert cold["converged"] and warm["converged"]
    assert warm["steps"] * 2 <= cold["steps"]
    assert abs(warm["drag"] - cold["drag"]) < 1e-3 * abs(cold["drag"])
