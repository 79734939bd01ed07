"""How much the reference's result depends on its dot-product partition.

Rdot_omp (mplapack/Rdot_omp.cpp) splits a dot product into one static chunk
per OpenMP thread and adds the partial sums, so the reference's iterates
depend on OMP_NUM_THREADS in the last bits.  On the dense random workload
(n=40, m=50) run to a relative gap of 1e-20, one chunk and 1024 chunks take
the same number of iterations to the same phase, with objectives that agree
to about 23 digits, not bit for bit: a bit-exact comparison with the
reference therefore has to fix the partition (the library's
sdpadd_set_dot_partition), which the GPU parity tests do.
"""
import mpmath
import pytest

from oracle_lib import Oracle  # noqa: F401  (fixture type)


def _mp(v):
    mpmath.mp.dps = 50
    return mpmath.mpf(v[0]) + mpmath.mpf(v[1])


def test_partition_spread_on_dense_random(oracle, tmp_path):
    from sdpadd import gen
    data = gen.random_dense(str(tmp_path / "c0.dat-s"))
    param = gen.param_file(str(tmp_path / "p.sdpa"), epsilon_star=1e-20, epsilon_dash=1e-20)
    one = oracle.solve(data, param, dot_threads=1)
    many = oracle.solve(data, param, dot_threads=1024)
    assert one["iterations"] == many["iterations"]
    assert one["phase"] == many["phase"] == "pdOPT"
    for k in ("objValPrimal", "objValDual"):
        a, b = _mp(one[k]), _mp(many[k])
        assert a != b, k                     # the partition is visible ...
        assert abs(a - b) / abs(a) < 1e-20, k  # ... but only far below the gap
    # same partition: bit-identical
    again = oracle.solve(data, param, dot_threads=1024)
    assert again["objValPrimal"] == many["objValPrimal"]
