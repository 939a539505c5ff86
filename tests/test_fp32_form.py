"""The FP32 precision form pinned rule by rule.

The form keeps parcel state, carrier velocity and drag in float and three things in
double: the cell test floor((x - origin) * rDelta), the residence velocity s0 with its
bounce corrections, and the deposits m (s0 - v') and their sums.  Inputs on meshes whose
cell widths are not powers of two are built so that taking any one of those rules in
float changes the result (checked here with the single-parcel statement in
fp32_form_spec.py, which follows asyncParallelTracking.jl:1271-1365 operation by
operation).  The oracle's FP32 form must then give the form's result exactly: parcel
state bit for bit, and the per-cell sums bit for bit (the serial oracle adds deposits in
parcel order, as the statement does).  The GPU library is held to the oracle on the same
inputs: state bit-exact, sources within 1e-12 (its FP64 additions run in another order).
"""
import math

import numpy as np
import pytest

import fp32_form_spec as spec
from oracle import st_oracle

F = np.float32
MU, RHO, RDR = 1e-3, 1000.0, 1.0


def face_floats(n, length, per_face=8):
    """Float coordinates next to the faces k * length / n for which the double and the
    float cell tests give different cells."""
    rd = 1.0 / (length / n)
    out = []
    for k in range(1, n):
        x = F(k * length / n)
        cand = [x]
        lo = hi = x
        for _ in range(per_face):
            lo = np.nextafter(lo, F(-1))
            hi = np.nextafter(hi, F(2 * length))
            cand += [lo, hi]
        for c in cand:
            if math.floor(float(c) * rd) != math.floor(float(F(F(c) * F(rd)))):
                out.append(c)
    return out


def family(name, seed=19891):
    """A mesh, a carrier field and a cloud of parcels (arrays of float32)."""
    rng = np.random.default_rng(seed)
    if name == "face":
        # 10 x 12 x 20 cells on the unit cube: every coordinate sits where the two cell
        # tests disagree, the parcels start at rest in a slow carrier, so most stay put
        n, e, dt, count = (10, 12, 20), (1.0, 1.0, 1.0), 1e-4, 400
        pools = [face_floats(n[a], e[a]) for a in range(3)]
        pos = [np.array(rng.choice(pools[a], count), dtype=F) for a in range(3)]
        vel = [np.zeros(count, F) for _ in range(3)]
        U = (rng.uniform(-1, 1, (n[0] * n[1] * n[2], 3)) * 1e-3).astype(F)
        dia = (10.0 ** rng.uniform(-5.3, -4.3, count)).astype(F)
    else:
        # 3^3 cells of 1/3 cm: carrier speeds and parcel velocities over four decades,
        # parcels from tracers to ballistic, frequent bounces and cell changes
        n, e, dt, count = (3, 3, 3), (0.01, 0.01, 0.01), 2e-3, 2000
        U = (rng.uniform(-1, 1, (27, 3)) * 10.0 ** rng.integers(-4, 1, (27, 1))).astype(F)
        pos = [rng.uniform(0, e[a], count).astype(F) for a in range(3)]
        vel = [(rng.uniform(-1, 1, count) * 10.0 ** rng.integers(-4, 1, count)).astype(F) for _ in range(3)]
        dia = (10.0 ** rng.uniform(-5.5, -4.0, count)).astype(F)
    mesh = spec.Mesh(n, (0.0, 0.0, 0.0), e)
    return {"n": n, "e": e, "dt": dt, "nsub": 10, "U": U, "d": dia, "pos": pos, "vel": vel, "mesh": mesh}


FAMILIES = {"face": "face", "s0": "mixed", "source": "mixed"}


def run_spec(fam, **rule):
    src = {}
    xs, vs = [], []
    for i in range(fam["d"].size):
        x, v, _ = spec.step(fam["mesh"], fam["U"], fam["d"][i], [p[i] for p in fam["pos"]],
                            [q[i] for q in fam["vel"]], fam["dt"], fam["nsub"], MU, RHO, RDR, into=src, **rule)
        xs.append(x)
        vs.append(v)
    return np.array(xs, F), np.array(vs, F), src


def per_parcel_differs(fam, rule):
    """Parcels whose own result (state or deposits) changes when `rule` is float."""
    k = 0
    for i in range(fam["d"].size):
        args = (fam["mesh"], fam["U"], fam["d"][i], [p[i] for p in fam["pos"]], [q[i] for q in fam["vel"]],
                fam["dt"], fam["nsub"], MU, RHO, RDR)
        a = spec.step(*args)
        b = spec.step(*args, **{rule: "float"})
        same_src = a[2].keys() == b[2].keys() and all(
            [float(t) for t in a[2][c]] == [float(t) for t in b[2][c]] for c in a[2])
        if not (a[0] == b[0] and a[1] == b[1] and same_src):
            k += 1
    return k


def run_oracle(fam):
    st = [fam["d"].copy()] + [p.copy() for p in fam["pos"]] + [q.copy() for q in fam["vel"]]
    src, _ = st_oracle.evolve(st, fam["U"], fam["n"], (0.0, 0.0, 0.0), fam["e"], fam["dt"], nsub=fam["nsub"],
                              precision=st_oracle.FP32, boundary=st_oracle.WALL, mu_c=MU, rho_d=RHO,
                              rho_d_by_rho_c=RDR)
    return st, src


def dense(src, ncell):
    out = np.zeros((ncell, 3))
    for c, v in src.items():
        out[c] = [float(t) for t in v]
    return out


@pytest.mark.parametrize("rule", list(FAMILIES))
def test_inputs_separate_the_rule(rule):
    """Taking the rule in float changes the outcome of many parcels of its family."""
    fam = family(FAMILIES[rule])
    k = per_parcel_differs(fam, rule)
    assert k >= 20, f"only {k} parcels separate the {rule} rule"


@pytest.mark.parametrize("name", ["face", "mixed"])
def test_oracle_fp32_form_follows_the_rules(name):
    fam = family(name)
    xs, vs, src = run_spec(fam)
    st, osrc = run_oracle(fam)
    ncell = fam["n"][0] * fam["n"][1] * fam["n"][2]
    np.testing.assert_array_equal(np.stack(st[1:4], 1), xs)
    np.testing.assert_array_equal(np.stack(st[4:7], 1), vs)
    np.testing.assert_array_equal(osrc, dense(src, ncell))
    # and the alternatives the family is built for are told apart on the whole cloud
    for rule in {"face": ("face", "source"), "mixed": ("s0", "source")}[name]:
        ax, av, asrc = run_spec(fam, **{rule: "float"})
        same = np.array_equal(ax, xs) and np.array_equal(av, vs) and np.array_equal(dense(asrc, ncell), osrc)
        assert not same, f"{name}: the {rule} rule taken in float gives the same cloud"


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["face", "mixed"])
@pytest.mark.parametrize("method", [0, 1])
def test_gpu_fp32_form_on_separating_inputs(ctx, name, method):
    import scaletrack as stk

    fam = family(name)
    st, osrc = run_oracle(fam)
    field = stk.TwoWayEulerian(ctx, stk.Mesh(fam["n"], (0.0, 0.0, 0.0), fam["e"]), stk.FP32)
    field.upload_U(fam["U"])
    field.zero_source()
    chunk = stk.Chunk(ctx, fam["d"].size, MU, RHO, RDR, stk.FP32)
    chunk.set_state(fam["d"], *fam["pos"], *fam["vel"])
    chunk.upload()
    p = stk.params(boundary=stk.WALL, method=method, brick=(2, 2, 2) if name == "mixed" else (4, 4, 4))
    stk.evolve_device(ctx, chunk, field, fam["dt"], fam["nsub"], p)
    got = chunk.download()
    gsrc = field.download_source()
    chunk.close()
    field.close()
    for a, b in zip(got, st):
        np.testing.assert_array_equal(a, b)
    err = np.abs(gsrc - osrc).max() / np.abs(osrc).max()
    assert err <= 1e-12, err
