"""SolverFactory for the device path: the reference's INI solver sections mapped onto the fused
solver's parameters (kernel/solver/solver_factory.hpp:376-640), and rejection of trees the fused
path does not implement.  CPU only (no solve); the GPU test compares a factory-built solve with
the directly constructed solver."""
import pytest

from feat3_amd import factory as fa

POISSON = """
# PCG + MG V(4,4), Richardson(Jacobi 0.5), exact coarse solve (applications/poisson_dirichlet.cpp)
[linsolver]
type = pcg
max_iter = 1000
tol_rel = 1e-8
plot_mode = iter
precon = mgv

[mgv]
type = mg
hierarchy = h1
lvl_min = -1
lvl_max = 0
cycle = v
memorytype = hip
datatype = float

[h1]
type = hierarchy
smoother = rich
coarse = coarse

[rich]
type = richardson
max_iter = 4
min_iter = 4
precon = jacobi

[jacobi]
type = jacobi
omega = 0.5

[coarse]
type = umfpack
"""


def test_read_ini_rules():
    pm = fa.read_ini("top = 1\n[a]\n x =  2 # c\n# only a comment\n\n[b]\ny=hello world\n[a]\nx = 3\n")
    assert pm[""] == {"top": "1"}
    assert pm["a"] == {"x": "3"}
    assert pm["b"] == {"y": "hello world"}
    with pytest.raises(fa.FactoryError):
        fa.read_ini("[a]\nnot a key value line\n")


def test_poisson_tree_maps_onto_the_fused_solver():
    kw = fa.solver_kwargs(fa.read_ini(POISSON))
    assert kw == dict(tol_rel=1e-8, max_iter=1000, inner="f32", omega_jacobi=0.5, omega_richardson=1.0, pre=4, post=4)


@pytest.mark.parametrize("edit,expect", [
    (("[mgv]", "datatype = float", "datatype = double"), dict(inner="f64")),
    (("[mgv]", "datatype = float", "datatype = float\nvaluetype = half"), dict(inner="f16")),
    (("[rich]", "max_iter = 4\nmin_iter = 4", "max_iter = 2\nmin_iter = 2\nomega = 0.9"),
     dict(pre=2, post=2, omega_richardson=0.9)),
    (("[linsolver]", "precon = mgv", "precon = none"), dict(precond="none")),
    (("[mgv]", "cycle = v", "cycle = V\nformat = banded"), dict(fmt="banded")),
    (("[mgv]", "cycle = v", "cycle = F"), dict(cycle="f")),
    (("[mgv]", "cycle = v", "cycle = w"), dict(cycle="w")),
    (("[h1]", "smoother = rich", "smoother = jacobi"), dict(pre=1, post=1, omega_jacobi=0.5, omega_richardson=1.0)),
    (("[coarse]", "type = umfpack", "type = cusolver"), dict()),
])
def test_variants(edit, expect):
    _, old, new = edit
    kw = fa.solver_kwargs(fa.read_ini(POISSON.replace(old, new)))
    for k, v in expect.items():
        assert kw[k] == v, (k, kw)


@pytest.mark.parametrize("old,new,msg", [
    ("type = pcg", "type = gropppcg", "linsolver.type"),
    ("cycle = v", "cycle = x", "mgv.cycle"),
    ("cycle = v", "", "mgv: mg section without cycle key"),
    ("type = umfpack", "type = pcg\nprecon = jacobi", "coarse.type"),
    ("max_iter = 4\nmin_iter = 4", "max_iter = 4", "rich"),
    ("datatype = float", "datatype = float\nmemorytype = cuda", "mgv.memorytype"),
    ("lvl_min = -1", "lvl_min = 2", "mgv"),
    ("hierarchy = h1", "hierarchy = h2", "h2"),
    ("precon = jacobi", "precon = none", "rich.precon"),
    ("tol_rel = 1e-8", "tol_rel = 1e-8\nstag_rate = 0.9", "linsolver.stag_rate"),
    ("max_iter = 1000", "max_iter = lots", "linsolver.max_iter"),
])
def test_unsupported_trees_are_rejected(old, new, msg):
    with pytest.raises(fa.FactoryError, match=msg.replace(".", r"\.")):
        fa.solver_kwargs(fa.read_ini(POISSON.replace(old, new, 1)))


def test_reference_example_is_rejected_with_reasons():
    """data/ini/solver_example.ini (gropppcg, F-cycle, Schwarz smoother, PCG coarse solver) is
    outside this path; the error names the first unsupported section."""
    ref_like = POISSON.replace("type = pcg", "type = gropppcg", 1)
    with pytest.raises(fa.FactoryError, match="gropppcg"):
        fa.solver_kwargs(fa.read_ini(ref_like))
