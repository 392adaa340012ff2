"""Where two runs of the same timestep may take different Picard counts.

The reference's Picard loop (run_nonlinear_matrixfree_split_mp.py:158-168)
stops when the residual falls below snes_atol, or when it rose (the last
iterate is then reverted).  Two runs that differ only by rounding (the GPU and
the C restatement sum in different orders) can part in two places, and only
there:

* threshold tie: at the iteration where they part, the residuals of both
  sides lie within `band` of snes_atol, the side that stopped below it and the
  side that went on at or above it;
* floor tie: both sides have stopped contracting (each side's residual at
  the parting iteration is at least `stall` times its previous one: the
  nonlinear iteration sits at its rounding floor), and the two residuals agree
  within `agree`; the rise that triggers the reference's revert is then
  rounding noise.

Any other parting is a real difference and fails the test."""


def parting_kind(rc, rg, atol, band=0.01, stall=0.9, agree=0.05):
    """rc, rg: residual after every Picard iteration of each side.  Returns
    None when both took the same count, else 'threshold' or 'floor'; raises
    AssertionError for any other parting."""
    kc, kg = len(rc), len(rg)
    if kc == kg:
        return None
    q = min(kc, kg) - 1
    stopped, went_on = (rc[q], rg[q]) if kc < kg else (rg[q], rc[q])
    if (1 - band) * atol <= stopped < atol <= went_on <= (1 + band) * atol:
        return 'threshold'
    if q >= 1:
        stalled = rc[q] >= stall * rc[q - 1] and rg[q] >= stall * rg[q - 1]
        if stalled and abs(rc[q] - rg[q]) <= agree * max(rc[q], rg[q]):
            return 'floor'
    raise AssertionError("Picard counts differ (%d vs %d) away from a tie: %s / %s" % (kc, kg, rc, rg))
