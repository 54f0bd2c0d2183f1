"""The tuner's result -> solver keyword mapping (CPU): launch-shape keys become solver options,
the format keys stay keyword arguments, a transfer rows-per-block of 0 is kept (it selects the
derived grid instead of the solver's default)."""
from feat3_amd import tune


def test_solver_kwargs_maps_launch_keys_to_options():
    kw = tune.solver_kwargs({"config": dict(fmt="sell", sell_C=64, sell_sigma=1, rows_per_block=256,
                                            transfer_rows_per_block=0)})
    assert kw == dict(fmt="sell", sell_C=64, sell_sigma=1,
                      options=dict(rows_per_block=256, transfer_rows_per_block=0))
    assert tune.solver_kwargs({"config": dict(fmt="banded", rows_per_block=0)}) == dict(fmt="banded")
    kw = tune.solver_kwargs({"config": dict(fmt="csr", lanes=8, options=dict(tail_max_n=0), transfer_rows_per_block=512)})
    assert kw == dict(fmt="csr", lanes=8, options=dict(tail_max_n=0, transfer_rows_per_block=512))
