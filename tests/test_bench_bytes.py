"""Algorithmic bytes per ADO-stage used by bench.py's roofline (CPU only).

The figures are those of the reference's stepping on the four measured
workloads (rkdp.h:108-146 without FSAL, rk4.h:40-65 fused), computed by hand:
  7-site chain, N=7, K=21, Dormand-Prince, complex128: 47/7*49*16 + 8*21 + 8
  N=64, K=16, Dormand-Prince, complex128:               47/7*4096*16 + 8*16 + 8
  sweep member N=8, K=4, RK4, complex128 / complex64:   5*64*16 + 40, 5*64*8 + 40
and the Dormand-Prince block count is checked by enumerating the stages.
"""
import bench


def test_per_unit_bytes_of_the_measured_workloads():
    assert bench.alg_bytes_per_ado_stage(7, 21, "rkdp") == 5440.0
    assert abs(bench.alg_bytes_per_ado_stage(64, 16, "rkdp") - 440163.43) < 0.01
    assert bench.alg_bytes_per_ado_stage(8, 4, "rk4") == 5160.0
    assert bench.alg_bytes_per_ado_stage(8, 4, "rk4", 8) == 2600.0


def test_dormand_prince_moves_47_blocks_per_attempt():
    # stage i (1..6): read its input, y and k_1..k_{i-1}; write k_i and the next input
    blocks = 0
    for i in range(1, 7):
        reads = {"in%d" % i, "y"} | {"k%d" % j for j in range(1, i)}
        writes = {"k%d" % i, "in%d" % (i + 1)}
        blocks += len(reads) + len(writes)
    # stage 7: read y_new (its input), write k_7, and the error vector sum E_j k_j
    # reads k_1..k_6
    blocks += len({"in7", "k7"} | {"k%d" % j for j in range(1, 7)})
    assert blocks == 47
    n2sc = 49 * 16
    assert bench.alg_bytes_per_ado_stage(7, 21, "rkdp") == blocks / 7 * n2sc + 8 * 21 + 8
