"""The tests that pin the oracle catch plausible mistakes in it.

The oracle is built again with one planted mistake at a time (oracle.c,
OR_PLANT; `make -C oracle plant`), and the checks that pin the affected
function run against that build: each must now fail.  A check that still
passed would be retyping the oracle rather than pinning it.

  1  BDPT connection: 1/d^2 dropped from G            (bdpt.cl:603-613)
  2  BDPT connection: weight w_i (1 - ks_i) -> w_i     (bdpt.cl:635)
  3  BDPT eye weights: ks floor 0.1 -> 0               (bdpt.cl:592)
  4  light vertex 0: r1 / r2 not swapped on the edges  (bdpt.cl:443-444)
  5  light vertex 0: contrib ke instead of ke * A      (bdpt.cl:452-453)
  6  light vertex 1: cos / pdf factor dropped          (bdpt.cl:467-468)
  7  BDPT connection: light-side cosine dropped        (bdpt.cl:613)
  8  BDPT connection: light-side BRDF dropped          (bdpt.cl:620-622)
  9  NEE: distance^2 dropped from the light pdf        (udpt.cl:662-665)
 10  MIS: balance instead of power heuristic           (udpt.cl:1149-1151)
 11  glossy BRDF: normaliser (n+1) instead of (n+2)    (udpt.cl:626)

Runs on the CPU only (gcc builds the variants).
"""
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

import test_oracle_bdpt as TB
import test_oracle_math as TM
from oracle import pyoracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PLANTS = {
    1: [TB.check_connection_sum, lambda: TB.check_two_plane_quadrature(1)],
    2: [TB.check_connection_sum, lambda: TB.check_two_plane_quadrature(1)],
    3: [TB.check_connection_sum, lambda: TB.check_two_plane_quadrature(1), TM.test_bdpt_furnace_closed_form],
    4: [TB.check_light_subpath],
    5: [TB.check_light_subpath, lambda: TB.check_two_plane_quadrature(1)],
    6: [TB.check_light_subpath, lambda: TB.check_two_plane_quadrature(1)],
    7: [TB.check_connection_sum, lambda: TB.check_two_plane_quadrature(1)],
    8: [TB.check_connection_sum, lambda: TB.check_two_plane_quadrature(1)],
    9: [TM.test_light_sampling_estimates_irradiance, lambda: TM.test_direct_light_converges_to_irradiance(0, 0)],
    10: [lambda: TM.test_direct_light_converges_to_irradiance(1, 0)],
    11: [TM.test_glossy_brdf_normalisation],
}


@pytest.fixture(scope="module")
def variants(tmp_path_factory):
    d = tmp_path_factory.mktemp("planted")

    def build(k):
        out = str(d / f"liboracle_plant{k}.so")
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "plant", f"K={k}", f"PLANT_OUT={out}"],
                       check=True, capture_output=True)
        return k, out

    with ThreadPoolExecutor(4) as ex:
        return dict(ex.map(build, PLANTS))


def test_default_build_passes_every_check():
    for checks in PLANTS.values():
        for chk in checks:
            chk()


@pytest.mark.parametrize("plant", sorted(PLANTS))
def test_planted_mistake_is_caught(plant, variants, monkeypatch):
    with O.variant(variants[plant]) as vl:
        monkeypatch.setattr(TM, "L", vl)      # test_oracle_math binds the library at import
        for chk in PLANTS[plant]:
            with pytest.raises(AssertionError):
                chk()
