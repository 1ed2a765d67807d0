"""Drop-in API end to end on the GPU against the reference's own golden solve
results (tests/_mpi_solve_reference.json of the reference): RlxPre (device
assembly, FP32 storage as Radia stores it), RlxAuto (device solver), Fld.

The golden values were produced by the reference's CPU path at its stopping
misfit (3.6e-6 / 8.0e-6 T); converging further moves B by 3.4e-9 (small) and
1.5e-6 (mixed) of the field scale, so the bounds below are 1e-7 and 5e-6."""
import json
import os

import numpy as np
import pytest

from radiacuda_amd import rad
from tests import golden_models

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("kind,tol", [("small", 1e-7), ("mixed", 5e-6)])
def test_golden_solve_through_dropin_api(kind, tol):
    print("case: golden", kind, flush=True)
    gold = json.load(open(os.path.join(HERE, "golden", "mpi_solve_reference.json")))[kind]
    ref = np.array(gold["cpu_asm_cpu_relax"]["b"])
    top, iron, coil = golden_models.build(kind)
    k = rad.RlxPre(top)
    res = rad.RlxAuto(k, 1e-9, 5000)
    print("relax", res, flush=True)
    B = np.array(rad.Fld(top, "b", golden_models.POINTS))
    err = np.abs(B - ref).max() / np.abs(ref).max()
    print("err vs golden %.3e" % err, flush=True)
    assert res[0] <= 1e-9
    assert err <= tol
    assert rad.UtiFldLastBackend() == "gpu" and rad.UtiAsmLastBackend() == "gpu"


def test_large_golden_recmag_model():
    """The reference's "large" golden (tests/test_mpi_solve.py:37-46): 1,100
    random RecMag blocks, chi = 20,000 iron, a racetrack coil.  Its
    cpu_asm_cpu_relax values are the method-4 state after the 3,000-sweep cap
    (misfit 0.174), which the oracle reproduces (tests/test_oracle.py).  Here:
    the device assembly against the oracle (every 3x3 block <= 1e-10), the
    device field and coil at that method-4 state against the reference's B
    (the reference test's 1e-6), and the device solver through the drop-in
    API to prec 1e-5 (where method 4 stopped short) against the FP64 fixed
    point of the device matrix."""
    from oracle import oracle as OR
    from radiacuda_amd import _native as N
    from radiacuda_amd import objects as O
    from tests import models
    print("case: golden large (1100 RecMag)", flush=True)
    gold = json.load(open(os.path.join(HERE, "golden", "mpi_solve_reference.json")))["large"]
    ref = np.array(gold["cpu_asm_cpu_relax"]["b"])
    B4, misfit4, M4 = golden_models.oracle_solve("large", return_state=True)
    assert abs(misfit4 - gold["cpu_asm_cpu_relax"]["misfit"]) <= 1e-2 * gold["cpu_asm_cpu_relax"]["misfit"]
    top, iron, coil = golden_models.build("large")
    P, items = models.packed(iron)
    n = len(items)
    assert n == 1100
    with N.Model(P) as m:
        Q = m.assemble(N.RH_FP64)
        Qd = Q.export()
        Qo, _mask = models.oracle_assemble(P)
        e_q = models.block_rel(Qd, Qo, n)
        cur = O.pack_currents(list(O.walk(rad._reg, coil)))
        pts = np.array(golden_models.POINTS)
        B = m.field(M4, pts)[0] + N.current_field(cur, pts)
        e_b = np.abs(B - ref).max() / np.abs(ref).max()
        print("large golden: device blocks vs oracle %.2e; device field at the method-4 state vs "
              "reference %.2e" % (e_q, e_b), flush=True)
        assert e_q <= 1e-10
        assert e_b <= 1e-6
        # the device solver through the drop-in API
        k = rad.RlxPre(top)
        res = rad.RlxAuto(k, 1e-5, 3000, 4)
        print("large golden: RlxAuto(1e-5, 3000) ->", res, flush=True)
        assert res[0] <= 1e-5
        mats = np.stack([o.mat.record(o.M) for (o, _k, _s) in items])
        Mg = np.stack([o.M for (o, _k, _s) in rad._reg[k].items])
        X = m.colloc()
        hext = N.current_field(cur, X)
        Mn = models.fp64_fixed_point(Qd, mats, hext, M0=Mg)
        dM = np.abs(Mg - Mn).max() / np.abs(Mn).max()
        print("large golden: device solve vs FP64 fixed point %.2e" % dM, flush=True)
        assert dM <= 1e-3
        Q.close()
