"""The "#define NAME value" parameter reader and the Rayleigh-number map (host code)."""
import math
from pathlib import Path

import pytest

from lagrangianrbnudging_amd import rbn

REF = Path("/root/reference")

TRUTH_BLOCK = """
#include <stdio.h>
#define REAL double
#define cs2 (1./3.)
#define cs1 sqrt((cs2))
#define AMIROOT (!me)
/////////INPUT VALUES/////////
#define nx         512
#define ny         256   // vertical
#define scratch    1
#define nstart     0
#define nsteps     20000
#define nout       1000
#define noutconfig 20
#define rhom1      1.0
#define relax      0.53
#define relaxg     (0.5+0.03)
#define Tup        0.0
#define Tdown      1.0
#define gravity    (-0.00002)
/* a block comment
#define nx 3
*/
#define Particle   1
#define Nprtcl 20*20
#define taulag 1500.0
#define onebytaulag (1/taulag)
#define half (1/2)
"""


def test_parse_truth_block():
    p = rbn.parse_params_text(TRUTH_BLOCK)
    assert (p.nx, p.ny, p.scratch, p.nsteps, p.nout, p.noutconfig) == (512, 256, 1, 20000, 1000, 20)
    assert p.relax == 0.53 and p.relaxg == 0.5 + 0.03
    assert p.gravity == -0.00002
    assert p.nprtcl == 400
    assert p.taulag == 1500.0
    assert p.variant == rbn.RBN_VARIANT_TRAJECTORIES


def test_c_integer_division_rules():
    p = rbn.parse_params_text("#define nx (7/2)\n#define ny 9/2*2\n#define relax (1/2.)\n")
    assert (p.nx, p.ny, p.relax) == (3, 8, 0.5)


def test_nudging_block_selects_variant():
    p = rbn.parse_params_text("#define alpha 0.1\n#define ntempread 5\n#define nx 64\n")
    assert p.variant == rbn.RBN_VARIANT_NUDGING and p.ntempread == 5 and p.alpha == 0.1
    p = rbn.parse_params_text("#define alpha 0.1\n#define rbn_variant 0\n")
    assert p.variant == rbn.RBN_VARIANT_TRAJECTORIES


def test_bad_parameter_value_is_a_parse_error():
    with pytest.raises(rbn.RbnError) as e:
        rbn.parse_params_text("#define nx foo(3)\n")
    assert e.value.code == -3


def test_missing_file():
    with pytest.raises(rbn.RbnError) as e:
        rbn.parse_params("/nonexistent/params.h")
    assert e.value.code == -2


def test_rayleigh_matches_media_formula():
    # media(): Ra = -(gravity/rhom1) L^3 dT / (nu kappa), nu = cs2 (relax - 1/2)
    p = rbn.default_params()
    nu = (1 / 3) * (p.relax - 0.5)
    ra = -(p.gravity / p.rhom1) * (p.ny - 1) ** 3 * (p.Tdown - p.Tup) / (nu * nu)
    assert math.isclose(rbn.rayleigh(p), ra, rel_tol=1e-14)
    assert 1.0e9 < rbn.rayleigh(p) < 1.1e9  # the reference's 1680x840 set-up


@pytest.mark.parametrize("Ra,Pr", [(1e6, 1.0), (1e8, 1.0), (1e7, 0.7)])
def test_from_ra_inverts_media(Ra, Pr):
    p = rbn.default_params(nx=512, ny=256, Tup=0.0, Tdown=1.0)
    q = rbn.params_from_ra(p, Ra, Pr, 0.05)
    assert math.isclose(rbn.rayleigh(q), Ra, rel_tol=1e-12)
    nu, kappa = (q.relax - 0.5) / 3, (q.relaxg - 0.5) / 3
    assert math.isclose(nu / kappa, Pr, rel_tol=1e-12)
    assert math.isclose(math.sqrt(-q.gravity * (q.ny - 1)), 0.05, rel_tol=1e-12)


@pytest.mark.skipif(not REF.exists(), reason="reference sources not mounted")
def test_reads_the_reference_define_blocks():
    p = rbn.parse_params(str(REF / "Trajectories/code/temppart_dump.c"))
    p = rbn.parse_params(str(REF / "Trajectories/code/mympi2.h"), p)
    d = rbn.default_params()
    for k in ("nx", "ny", "nsteps", "nout", "noutconfig", "relax", "relaxg", "Tup", "Tdown", "gravity",
              "nprtcl", "tracer", "heavyprtcl", "taulag", "beta", "prtcl_writefull"):
        assert getattr(p, k) == getattr(d, k), k
    assert p.variant == rbn.RBN_VARIANT_TRAJECTORIES
    q = rbn.parse_params(str(REF / "Nudging/code/tempvel_nudge.c"))
    q = rbn.parse_params(str(REF / "Nudging/code/mympi2.h"), q)
    assert q.variant == rbn.RBN_VARIANT_NUDGING
    assert (q.ntempread, q.alpha, q.nprtcl, q.scratch, q.noutconfig) == (250, 0.01, 1600, 1, 20000)
