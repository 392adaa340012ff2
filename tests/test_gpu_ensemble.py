"""The batched uneven-grid ensemble (256 instances from
make_instances(256, seed=0): nx in {64,128,256,512}, nv in {128,...,1024},
k ~ U[0.2, 0.6], amplitude ~ U[0.005, 0.5], Landau or bump-on-tail, J4,
dt = 0.05) through the bucketed batched driver with captured-graph replay
and autotuned launch shapes, checked instance by instance against the C
restatement of the reference's timestep (oracle/c), which runs each instance
alone as the reference's run script does
(run_nonlinear_matrixfree_split_mp.py:111-174).

Bars (written in the test), after each of 2 timesteps and for every
instance: the same Picard count, and f within 1e-12 relative (max norm).
Where the two runs part at a threshold tie or a floor tie
(tests/picard_ties.py: the 512 x 1024 instances' residual has its rounding
floor at about snes_atol = 1e-10, where the reference's stop decision is
rounding noise), f must still agree within 1e-12 relative, the same bar;
the C run is then set to the GPU state and the comparison goes on.  At most 20 % of the 512 instance-steps may be such
ties.  Any other difference in Picard counts fails."""

import numpy as np
import pytest

from oracle import vlasov_oracle as O

pytestmark = pytest.mark.gpu

STEPS = 2
ATOL = 1e-10


def test_ensemble_of_256_matches_c_oracle(cuda, progress):
    from oracle import cbaseline
    from picard_ties import parting_kind, tie_f_rtol
    from vivlasov1d_amd.batch import BatchRun, make_instances
    specs = make_instances(256, seed=0)
    print("case ensemble 256 instances, %d buckets" % len({(s['nx'], s['nv']) for s in specs}), flush=True)
    br = BatchRun(specs, graph=True, autotune=True)
    where = {}
    for bi, b in enumerate(br.buckets):
        for q, s in enumerate(b.specs):
            where[s['index']] = (bi, q)
    states = [{s['index']: br.state(s['index'])[0] for s in specs}]
    recs = []
    for t in range(STEPS):
        recs.append(br.step())
        br.synchronize()
        states.append({s['index']: br.state(s['index'])[0] for s in specs})
        print("batched step", t + 1, "done", flush=True)
    ties, worst = 0, 0.0
    for n, s in enumerate(specs):
        if n % 32 == 0:
            progress("  ensemble: C restatement of instance %d of %d" % (n, len(specs)))
        bi, q = where[s['index']]
        g = O.Grid(s['nx'], s['nv'], s['L'], s['vmax'], ht=s['ht'])
        crun = cbaseline.CRun(g, states[0][s['index']], bracket='arakawaj4')
        for t in range(STEPS):
            ec = crun.step()
            rec = recs[t][bi]
            assert len([r for r in rec['residuals'][:, q] if r == r]) == int(rec['picard'][q])
            kind = parting_kind(ec['residuals'], [r for r in rec['residuals'][:, q] if r == r], ATOL)
            if kind is not None:
                ties += 1
                fc = crun.f()
                err = np.abs(states[t + 1][s['index']] - fc).max() / np.abs(fc).max()
                print(kind, "tie: instance", s['index'], "step", t + 1, "f error", err, flush=True)
                assert err <= tie_f_rtol, (s['index'], t, err)
                crun.set_f(states[t + 1][s['index']])
                continue
            fc = crun.f()
            err = np.abs(states[t + 1][s['index']] - fc).max() / np.abs(fc).max()
            worst = max(worst, err)
            assert err <= 1e-12, (s['index'], s['nx'], s['nv'], t, err)
        crun.close()
    print("worst f error over the ensemble", worst, "threshold ties", ties, flush=True)
    assert ties <= 0.2 * len(specs) * STEPS, ties
