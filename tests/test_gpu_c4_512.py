"""The 512^3 scale-up configuration on the GPU: seeded random ellipsoids
(semi-axes 8-24 voxels, major axis within 15 degrees of the through-thickness
axis d2) with a random-walk tunnel of radius 3, cleaned of disconnected solid
and electrolyte by the GPU labeller (the RM_UCNCT_3D step), then coupled
electrochemical steps (SBMES_3D_MPI_CH_Gr_CONJG_VD_T05_01_BQ_hx3.f90:397-840)
with the benchmark's fast path (FP32 CG preconditioned by a multigrid V-cycle,
FP64 reliable updates, the learned warm start),
the CnE solve resolving each step's change (as the 256^3 benchmark).
Every solve of every step reaches relres 1e-8 with an FP64 true residual, the
fields stay finite, Xrf changes only by the step's reaction, and the device
footprint is reported (and stays well inside one GPU's 288 GB)."""
import numpy as np
import pytest

import elsim_amd as E
from elsim_amd import nn as NN
from elsim_amd import scaleup
from elsim_amd._lib import F_CNE, F_CNP, F_PHE, F_PHP

pytestmark = pytest.mark.gpu
N = 512
STEPS = 3


def test_c4_512_coupled_steps(gpu):
    base, total = scaleup.mem_used()
    sdf, ns, frac = E.ellipsoid_pack(N, 3, 8.0, 24.0, cone_deg=15.0, target=0.5, overlap=0.3, dcut=12.0)
    sdf = E.carve_tunnel(sdf, 3, radius=3.0)
    sdf, cstats = scaleup.clean(sdf, 12.0)
    s = E.Sbm(sdf, -sdf, E.synthetic_params(ws_mode=E.WS_LEARNED, cne_rule=1, ws_order=2), E.tables())
    s.set(F_CNP, E.initial_cnp(N, 3))
    net = NN.WarmStartNet.load(NN.DEFAULT_WEIGHTS)
    s.set_warm_start(net)
    xrf0 = s.scalars()["xrf"]
    for k in range(STEPS):
        st = s.step(E.MODE_MIXED_MG)
        print(f"step {k}: phP {st.it_php} phE {st.it_phe} CnE {st.it_cne} outer {st.n_outer} "
              f"relres {st.relres_php:.1e} {st.relres_phe:.1e} {st.seconds:.2f} s", flush=True)
        assert st.status == 0, st.as_dict()
        assert st.relres_php <= 1e-8 and st.relres_phe <= 1e-8
        if k > 0:
            assert st.relres_cne <= 1e-8
    used, _ = scaleup.mem_used()
    footprint = used - base
    print(f"{N}^3: {ns} ellipsoids, solid {frac:.3f}, clean-up {cstats}; device memory {footprint / 2**30:.1f} GiB "
          f"of {total / 2**30:.0f} GiB", flush=True)
    assert footprint < 0.5 * total
    for f in (F_CNP, F_CNE, F_PHP, F_PHE):
        assert np.isfinite(s.get(f)).all()
    assert abs(s.scalars()["xrf"] - xrf0) < 1e-3
    s.close()
    net.close()
