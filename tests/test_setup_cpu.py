"""Host setup (mesh, body mask, solve boxes, EnteredData.txt format) against
the reference's own output files for the golden cases."""
import os

import numpy as np
import pytest

from golden_cases import ALL_CASES, input_dir, ref_dir, text_mesh

MESH_FILES = ["x_f", "y_f", "hx", "hy", "x_v", "y_v"]


@pytest.mark.parametrize("name", ALL_CASES)
def test_mesh_and_mask_match_reference(sts, name):
    c = sts.Case(input_dir(name))
    for f in MESH_FILES:
        ref = open(os.path.join(ref_dir(name), f + ".txt")).read()
        assert text_mesh(getattr(c, f)) == ref, f"{name}: {f}.txt differs from the reference's"
    vol = np.fromfile(os.path.join(ref_dir(name), "vol_inf_fb.bin"), dtype=np.uint32)
    assert np.array_equal(vol, c.vol), f"{name}: body mask differs"


@pytest.mark.parametrize("name", ALL_CASES)
def test_config_rewrite_format(sts, name, tmp_path):
    """EnteredData.txt as the reference rewrites it: identical on every line
    the run does not change (elapsed hours, finish flag and the inflow
    pressure the run updates are excluded)."""
    c = sts.Case(input_dir(name))
    out = tmp_path / "EnteredData.txt"
    c.write_config(str(out))
    mine = out.read_text().split("\n")
    ref = open(os.path.join(ref_dir(name), "EnteredData.txt")).read().split("\n")
    assert len(mine) == len(ref)
    varying = {44, 45, 77}  # max_time_for_solve, solve_is_finished, p_BC_xb
    if c.value("use_interpolation_for_u_BC_by_time"):
        varying |= {1, 71}  # u_gas_nd, u_given_xb follow the inflow ramp
    diff = [k + 1 for k in range(len(ref)) if mine[k] != ref[k] and (k + 1) not in varying]
    assert not diff, f"{name}: lines {diff} differ: {[(mine[k - 1], ref[k - 1]) for k in diff[:3]]}"


def test_boxes_follow_define_area_for_solving(sts):
    """S:6009-6267: pressure-driven and given-velocity boxes."""
    c = sts.Case(input_dir("shipped_coarse"))
    Nx, Ny = c.Nx, c.Ny
    assert c.boxes == {"Nx_u": (1, Nx - 1), "Ny_u": (1, Ny - 2), "Nx_v": (1, Nx - 2), "Ny_v": (1, Ny - 1),
                       "Nx_o": (1, Nx - 2), "Ny_o": (1, Ny - 2)}
    c = sts.Case(input_dir("square"))
    Nx, Ny = c.Nx, c.Ny
    assert c.boxes["Nx_u"] == (2, Nx - 1) and c.boxes["Nx_o"] == (1, Nx - 2)


def test_type3_constants(sts):
    """S:2057-2071: mode 3 overwrites Ma, Pr, gamma1, c_mu and derives cT, cT1, cT2."""
    c = sts.Case(input_dir("shipped_coarse"))
    Kn = c.value("Kn")
    g, Ma, Pr = 5.0 / 3.0, np.sqrt(6.0 / 5.0), 2.0 / 3.0
    c_mu = (5.0 / 16.0) * np.sqrt(2.0 * np.pi / g)
    assert c.value("Ma") == Ma and c.value("gamma1") == g and c.value("c_mu") == c_mu
    assert c.value("cT") == Kn * 5.0 * np.sqrt(np.pi) / 16.0
    assert c.value("cT1") == g * c_mu * Kn / (Ma * Pr)


def test_missing_input_fails_loudly(sts, tmp_path):
    with pytest.raises(sts.StsError):
        sts.Case(str(tmp_path))
    (tmp_path / "EnteredData.txt").write_text("1\t//u_gas_nd\n")
    with pytest.raises(sts.StsError):
        sts.Case(str(tmp_path))
