"""What the trained warm-start network gains over the method without it.

The committed network (topopt.jl_amd/weights/warmstart_64x32x32.npz) was
trained on the GPU by the library's own BF16 kernels, on (x~, u) pairs from
plain SIMP runs of the 64 x 32 x 32 cantilever (volume fraction ~ U[0.25,
0.6], load node uniform on the free face; hold-out loss 0.0049 on u/rms(u)).
Here it is run on designs it did not see: three SIMP runs with fixed volume
fractions and load positions, each solved three ways (from zero, from the
previous displacement, and from the point of span{previous displacement,
network prediction} with the smallest residual), all to the same residual
tolerance ||f - K u|| <= 1e-8 ||f||, with Jacobi and with multigrid PCG.

Measured on MI355X: the network barely pays at this tolerance.  Its
prediction (about 7 % RMS error) starts CG with error in components PCG
resolves slowly: first solves from the prediction alone took 17 % (Jacobi) and
6 % (multigrid) more iterations than from zero; whole runs took 2 % more
(Jacobi) and 1.7 % fewer (multigrid) than with the previous displacement
alone.  The test asserts what must hold whatever the network's quality (PCG
still judges every solve, so the answers do not change), and reports the
iteration counts.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RUNS = [(0.30, 16, 16), (0.45, 4, 28), (0.55, 30, 3)]  # volfrac, load node j, k


@pytest.mark.parametrize("precision", ["mixed", "mixed_mg"])
def test_trained_network_versus_previous_iterate(T, gpu, precision):
    import topopt_jl_amd.nn as NN
    from topopt_jl_amd._lib import PRECISIONS
    nels = (64, 32, 32)
    assert os.path.exists(NN.DEFAULT_WEIGHTS)
    ctx, fixed = NN.make_context(nels)
    net = NN.WarmStartNet(nels, max_batch=1, fixed=fixed)
    meta = net.load(NN.DEFAULT_WEIGHTS)
    assert int(meta["pairs"]) == 2000
    totals = {"zero": 0, "previous": 0, "learned": 0}
    first = {"zero": 0, "learned": 0}
    for i, (vf, j, k) in enumerate(RUNS):
        spec = NN.SimpRunSpec(volfrac=vf, load_j=j, load_k=k, seed=10_000 + i)
        res = NN.compare_warm_starts(ctx, net, nels, spec, iters=10,
                                     precision=PRECISIONS[precision])
        # PCG to the same tolerance judges every solve: same designs, same answers
        np.testing.assert_allclose(res["learned"][1], res["previous"][1], rtol=1e-6)
        np.testing.assert_allclose(res["zero"][1], res["previous"][1], rtol=1e-6)
        for name in totals:
            its = res[name][0]
            assert np.all(its > 0) and np.all(its < 200_000)  # every solve converged
            totals[name] += int(its.sum())
        first["zero"] += int(res["zero"][0][0])
        first["learned"] += int(res["learned"][0][0])
        print(f"run {i} volfrac {vf} load ({j},{k}): first solve zero {res['zero'][0][0]} "
              f"learned {res['learned'][0][0]}; run totals " +
              ", ".join(f"{n} {int(res[n][0].sum())}" for n in totals))
    print(f"{precision}: first solves: zero {first['zero']} learned {first['learned']}; runs: " +
          ", ".join(f"{n} {v}" for n, v in totals.items()))
    # the previous-displacement warm start (the reference's initially_zero=false)
    # must beat starting every solve from zero
    assert totals["previous"] < totals["zero"]


def test_destroyed_network_detaches_from_solver(T, gpu):
    """A context keeps a pointer to its attached network; destroying the
    network detaches it, so a learned warm start afterwards is a state error
    (not a use of freed memory), and the other orders of destruction are safe."""
    import topopt_jl_amd.nn as NN
    nels = (16, 8, 8)  # multiples of 8: three 2x poolings
    prob = T.PointLoadCantilever(nels)
    s = T.FEASolver(T.HIPMatrixFreeSolver, prob, xmin=1e-9, penalty=T.PowerPenaltyFun(3.0))
    net = NN.WarmStartNet.for_problem(prob)
    s.attach_warm_start(net)
    res = T.simp_run(s, volfrac=0.3, rmin=1.5, iterations=2, warm_start="learned")
    assert res.converged
    net.close()
    with pytest.raises(T.TohipError, match="tohip_nn_attach"):
        T.simp_run(s, volfrac=0.3, rmin=1.5, iterations=1, warm_start="learned")
    # the plain paths still work on the same context
    assert T.simp_run(s, volfrac=0.3, rmin=1.5, iterations=1).converged
    # attach to a second context, destroy the contexts first, then the network
    net2 = NN.WarmStartNet.for_problem(prob)
    s2 = T.FEASolver(T.HIPMatrixFreeSolver, prob, xmin=1e-9, penalty=T.PowerPenaltyFun(3.0))
    s.attach_warm_start(net2)
    s2.attach_warm_start(net2)
    s.close()
    s2.close()
    net2.close()
