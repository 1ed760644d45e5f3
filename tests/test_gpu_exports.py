"""Entry points a driver links that the per-step path does not call itself,
against the compiled reference's values on the golden states:
md_calc_temperature_x (md_statistics_x.c:855-958) on every golden case, and
the Ewald energy splits md_coulomb_3d_ewald_energy_R_x / _R2_x / _NI_x
(md_statistics_x.c:398-495) through md_pimd_calc_pair_energy_x on the UEG
cases.  Bars: 1e-12 relative."""
import pytest

import pyxmd
from orc import nhc_from, read_golden

pytestmark = pytest.mark.gpu

CASES = ["trap16", "trap16_g3", "trap8_xi0", "trap8_xi05", "trap24_xi025", "ueg8_n2", "ueg8_n4_xi0", "ueg16_n8",
         "pgauss12"]


def system(lib, name):
    gold = read_golden()
    g = {k.split("/", 1)[1]: v for k, v in gold.items() if k.startswith(name + "/")}
    N, P, T, vi, kind, image, L, nsum, fgroup, dt = g["meta"]
    kind = ("trap", "ueg", "pgauss")[int(kind)]
    s = pyxmd.System(lib, kind, int(N), int(P), T, vi, L=L, nsum=int(nsum), fgroup=int(fgroup), positions=g["x0"])
    pyxmd.load_state(s, g["x0"], g["v0"], nhc_from(g["nhc0"]))
    return s, g


@pytest.mark.parametrize("name", CASES)
def test_temperature(name):
    lib = pyxmd.Lib()
    s, g = system(lib, name)
    st = lib.md_alloc_stats_x(1)
    st.contents.Ts[0] = 0.0
    assert lib.md_calc_temperature_x(s.sim, st) == 0
    assert st.contents.Ts[0] == pytest.approx(g["e_temperature"][0], rel=1e-12)


@pytest.mark.parametrize("name", ["ueg8_n2", "ueg8_n4_xi0", "ueg16_n8"])
def test_ewald_energy_splits(name):
    lib = pyxmd.Lib()
    s, g = system(lib, name)
    for q, cb in enumerate(["md_coulomb_3d_ewald_energy_R_x", "md_coulomb_3d_ewald_energy_R2_x",
                            "md_coulomb_3d_ewald_energy_NI_x"]):
        st = lib.md_alloc_stats_x(1)
        assert lib.md_pimd_calc_pair_energy_x(s.pimd, lib.cb[cb], st) == 0
        got, ref = st.contents.es[0], g["e_pair_R_R2_NI"][q]
        print(f"{name} {cb}: {got:.15g} vs {ref:.15g}")
        assert got == pytest.approx(ref, rel=1e-12, abs=1e-12)
