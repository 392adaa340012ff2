"""Where two runs of the same timestep may take different Picard counts.

The reference's Picard loop (run_nonlinear_matrixfree_split_mp.py:158-168)
stops when the residual falls below snes_atol, or when it rose (the last
iterate is then reverted).  Two runs that differ only by rounding (the GPU and
the C restatement sum in different orders) can part in two places, and only
there:

* threshold tie: at the iteration where they part, the side that stopped is
  below snes_atol, the side that went on at or above it, and the two
  residuals agree within `band`;
* floor tie: both sides have reached the rounding floor of the residual
  ||F(f)|| + ||A phi + (n - mean n) q||, where the stop decision (below
  snes_atol or not, rose or not) is rounding noise: at the parting iteration
  both residuals are within `near` times snes_atol and within a factor
  `ratio` of each other, or both have stopped contracting (each at least
  `stall` times its previous value) and agree within `agree`.  On the
  largest grids (512 x 1024) that floor lies at about snes_atol = 1e-10 and
  the two implementations' values there differ by up to tens of percent.

A parting anywhere else (a residual still on its contraction line, or the
two sides far apart) is a real difference.  A test that accepts a tie also
checks that the two states still agree to `tie_f_rtol`, the bar of the steps
where the counts agree (the extra Picard update at the floor moves f by
about 1e-15 relative), before it resynchronises them."""

tie_f_rtol = 1e-12


def parting_kind(rc, rg, atol, band=0.05, near=3.0, ratio=2.0, stall=0.9, agree=0.1):
    """rc, rg: residual after every Picard iteration of each side.  Returns
    None when both took the same count, else 'threshold' or 'floor'; raises
    AssertionError for any other parting."""
    kc, kg = len(rc), len(rg)
    if kc == kg:
        return None
    q = min(kc, kg) - 1
    a, b = rc[q], rg[q]
    stopped, went_on = (a, b) if kc < kg else (b, a)
    if stopped < atol <= went_on and went_on - stopped <= band * went_on:
        return 'threshold'
    if max(a, b) <= near * atol and max(a, b) <= ratio * min(a, b):
        return 'floor'
    if q >= 1 and a >= stall * rc[q - 1] and b >= stall * rg[q - 1] and abs(a - b) <= agree * max(a, b):
        return 'floor'
    raise AssertionError("Picard counts differ (%d vs %d) away from a tie: %s / %s" % (kc, kg, rc, rg))
