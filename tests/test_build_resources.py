"""Kernel resources recorded by the build (hipcc's kernel-resource-usage remarks, per translation
unit): the step kernels of the default scheme (MC + HLLC, FP32), and those the reference's own
scheme (Koren + isentropic) runs on, use no scratch memory, whose stores go to HBM on every plane
of a march, and the role-tiled kernels keep the occupancy and LDS their launch shapes are
designed for. CPU only."""
import pytest

from shaderous_amd import build

# mangled-name prefixes of the default scheme's FP32 step kernels: LIM 1 (MC), RS 0 (HLLC),
# ISEN false
NO_SCRATCH = {
    "k_fused_r (every role tile, boundary form and the flux-debug form)": "_ZN3shx9k_fused_rIfLi1ELi0ELb0E",
    "k_predict": "_ZN3shx9k_predictIfLi1ELb0E",
    "k_update": "_ZN3shx8k_updateIfLi1ELi0ELb0E",
    "k_zplanes": "_ZN3shx9k_zplanesIfLi1ELb0E",
    "k_ens_fused": "_ZN3shx11k_ens_fusedIfLi1ELi0ELb0E",
    "k_ens_rows": "_ZN3shx10k_ens_rowsILi1ELi0ELb0E",
    "k_amr_assemble": "_ZN3shx14k_amr_assembleIfLi1E",
    "k_amr_predict": "_ZN3shx13k_amr_predictIfLi1ELb0E",
    "k_amr_update": "_ZN3shx12k_amr_updateIfLi1ELi0ELb0E",
    "k_amr_shell (half steps)": "_ZN3shx11k_amr_shellIfLi1ELb0ELi1E",
}


# the reference's default scheme (Koren + isentropic reconstruction, HLLC; LIM 0, RS 0, ISEN
# true) in the shapes its runs take: the 32 x 16 role tile on a grid periodic in x and y (the
# autotuner's pick: 32 x 8 and 16 x 16 spill at 128 VGPRs with Koren) and the 2D row march
NO_SCRATCH_KOREN = {
    "k_fused_r Koren + isentropic 32 x 16": "_ZN3shx9k_fused_rIfLi0ELi0ELb1ELi32ELi16ELb0ELb0E",
    "k_ens_rows Koren + isentropic": "_ZN3shx10k_ens_rowsILi0ELi0ELb1E",
}


@pytest.fixture(scope="module")
def resources():
    res = build.kernel_resources()
    if not res:
        pytest.skip("no kernel resources recorded: build first (python -m shaderous_amd.build)")
    return res


def test_default_scheme_step_kernels_use_no_scratch(resources):
    for what, prefix in {**NO_SCRATCH, **NO_SCRATCH_KOREN}.items():
        found = [k for k in resources if k["kernel"].startswith(prefix)]
        assert found, f"no {what} kernel in the build"
        bad = [(k["kernel"], k.get("scratch")) for k in found if k.get("scratch", 0) != 0]
        assert not bad, f"{what} uses scratch memory: {bad}"


def test_role_tiles_keep_their_occupancy_and_lds(resources):
    # (TX, TY) -> (waves per SIMD, largest LDS per workgroup): two 7-wave workgroups per CU for
    # 32 x 8 and 16 x 16 (128 VGPRs, at most 80 KB each), one 12-wave one for 32 x 16
    want = {(32, 8): (4, 80 * 1024), (16, 16): (4, 80 * 1024), (32, 16): (3, 160 * 1024)}
    for (tx, ty), (occ, lds) in want.items():
        prefix = f"_ZN3shx9k_fused_rIfLi1ELi0ELb0ELi{tx}ELi{ty}E"
        found = [k for k in resources if k["kernel"].startswith(prefix)]
        assert found, (tx, ty)
        for k in found:
            assert k["occupancy"] >= occ and k["lds"] <= lds, (tx, ty, k)


def test_ensemble_default_tile_keeps_three_workgroups_per_cu(resources):
    # k_ens_fused 32 x 16 threads (28 x 12 cells, 8 waves, 51 KB of LDS) in its ensemble form
    # (instances from the device list): three workgroups per CU need six waves per SIMD (at most
    # 80 VGPRs); at five the ensemble step measured 7% slower
    prefix = "_ZN3shx11k_ens_fusedIfLi1ELi0ELb0ELi32ELi16ELb0E"
    found = [k for k in resources if k["kernel"].startswith(prefix)]
    assert found
    for k in found:
        assert k["occupancy"] >= 6 and k["lds"] <= 160 * 1024 // 3, k
