"""An end-to-end usage example written as a user of the Python host package would write it
(telegraph model, adaptive solve, then training records, device training and a
learned-adapter solve), with its results checked."""
import numpy as np
import pytest

import numcme_hip as nh

pytestmark = pytest.mark.gpu


def test_integration_example(gpu):
    S = np.array([[-1, 1, 0, 0], [1, -1, 0, 0], [0, 0, 1, -1]])  # telegraph
    props = [nh.propensity(lambda x, p: p[0] * x[0]), nh.propensity(lambda x, p: p[1] * x[1]),
             nh.propensity(lambda x, p: p[2] * x[1]), nh.propensity(lambda x, p: p[3] * x[2])]
    model = nh.CmeModel(S, props, [0.05, 0.1, 5.0, 1.0])
    alg = nh.AdaptiveFspSparse(ode_method=nh.KrylovExpv(m=30, tol=1e-12),
                               space_adapter=nh.SelectiveRStepAdapter(10, 10, True))
    out = nh.solve(model, nh.FspVectorSparse([[1, 0, 0]], [1.0]), (0.0, 120.0), alg,
                   saveat=np.arange(0.0, 121.0, 20.0), fsptol=1e-4)
    mass = [p.sum() + s.sum() for p, s in zip(out.p, out.sinks)]
    assert out.t[:7] == list(np.arange(0.0, 121.0, 20.0))
    # dropped states take their probability with them: the mass stays within the budget
    assert all(1.0 - 1e-4 <= m <= 1.0 + 1e-9 for m in mass)

    # learned adapter: records from plain runs, device training, learned solve
    rec = nh.TrainingRecords(capacity=200_000, per_adaptation=1024, label_mass=1e-10)
    nh.solve(model, nh.FspVectorSparse([[1, 0, 0]], [1.0]), (0.0, 60.0),
             nh.AdaptiveFspSparse(ode_method=nh.RK45(), space_adapter=nh.SelectiveRStepAdapter(2, 2, False)),
             saveat=[60.0], fsptol=1e-4, records=rec)
    assert len(rec) > 0
    net, loss = nh.train_adapter(*rec.arrays(), epochs=20)
    assert np.isfinite(loss)
    alg = nh.AdaptiveFspSparse(ode_method=nh.RK45(),
                               space_adapter=nh.LearnedAdapter(net, 2, 2, False, threshold=0.3))
    out = nh.solve(model, nh.FspVectorSparse([[1, 0, 0]], [1.0]), (0.0, 120.0), alg, saveat=[120.0], fsptol=1e-4)
    st = out.stats
    assert st["learned_adaptations"] + st["fallback_adaptations"] == st["adaptations"]
    assert out.sinks[-1].sum() <= 1e-4 + 1e-13
    assert abs(out.p[-1].sum() + out.sinks[-1].sum() - 1.0) <= 1e-9
