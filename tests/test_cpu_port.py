"""The CPU baseline port (oracle/cpu_port.c + scipy.fft) computes the oracle's
matvec (compute_Ax2 restated in oracle/pfft_ref.py) on the same operator."""
import numpy as np

from helpers import oracle_operator, rel, sphere_case


def test_port_matvec_equals_oracle():
    from oracle import cpu_port, pfft_ref
    c, e, p, k = sphere_case(n_cells=5)
    op = oracle_operator(n_cells=5)
    S4 = np.stack([op.S, op.Sx, op.Sy, op.Sz])
    port = cpu_port.PortOperator(c, e, k, S4, op.sp_mat, op.sp_ind + 1)
    rng = np.random.default_rng(3)
    x = rng.standard_normal(3 * e.shape[0]) + 1j * rng.standard_normal(3 * e.shape[0])
    y = port.matvec(x, 2)
    assert rel(y, pfft_ref.matvec(op, x)) < 1e-12
