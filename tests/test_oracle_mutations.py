"""The oracle pins catch plausible mistakes.

Each case plants one mistake in a copy of oracle/recon.py (a dropped term, a
flipped sign, swapped weights, a wrong constant), loads the copy as a module
and runs the pins of tests/test_oracle_pins.py that cover that code against
it: every listed pin must fail.  The unmodified oracle passes the same pins
(tests/test_oracle_pins.py).  Mistakes are planted on the CPU only.
"""
from __future__ import annotations

import types
from pathlib import Path

import pytest

from oracle import recon
from tests import test_oracle_pins as P

SRC = Path(recon.__file__).read_text()

# (name, text in oracle/recon.py, replacement, pins that must fail)
MUTATIONS = [
    ("cone weight without the v^2 term",
     "np.sqrt(sdd * sdd + U * U + V * V)", "np.sqrt(sdd * sdd + U * U)",
     [P.pin_cone_weight_is_ray_cosine, P.pin_fdk_of_z_invariant_cylinder]),
    ("curved weight in the flat form",
     "p = p * sdd * np.cos((T(-1) * U / sdd).astype(np.float64)).astype(dtype) * sdd / np.sqrt(sdd * sdd + V * V)",
     "p = (p * sdd * sdd) / np.sqrt(sdd * sdd + U * U + V * V)",
     [P.pin_cone_weight_is_ray_cosine]),
    ("curved back-projection with the atan2 sign flipped",
     "np.arctan2(yr.astype(np.float64)", "np.arctan2(-yr.astype(np.float64)",
     [P.pin_curved_fbp_of_offcentre_disc, lambda R: P.pin_fdk_of_z_invariant_cylinder(R, True)]),
    ("cone forward projection with the z weights swapped",
     "((T(1) - zw) * lo + zw * hi)", "(zw * lo + (T(1) - zw) * hi)",
     [P.pin_cone_fp_is_trilinear_line_sum, P.pin_cone_fp_of_ball_is_chord]),
    ("noise without the sqrt(N) scale",
     "g.astype(np.float32) * np.sqrt(t)", "g.astype(np.float32)",
     [P.pin_noise_statistics]),
    ("noise clamp at another count",
     "np.where(t <= 0, np.float32(1e-6), t)", "np.where(t <= 0, np.float32(1e-3), t)",
     [P.pin_noise_statistics]),
    ("IR pixel_count_ratio set to 1",
     "ratio = sgm.size / x0.size", "ratio = 1.0",
     [P.pin_ir_solve_is_cg_on_the_normal_equations, P.pin_piccs_is_cg_on_its_normal_equations]),
    ("PICCS without the prior term on the right-hand side",
     "b = bpb + (c2 * dtwd_(np.asarray(prior, np.float64), wp) if prior is not None else 0.0)",
     "b = bpb",
     [P.pin_piccs_is_cg_on_its_normal_equations]),
    ("PICCS prior weights from x instead of x - prior",
     "wp = tvw(x - prior)", "wp = tvw(x)",
     [P.pin_piccs_is_cg_on_its_normal_equations]),
    ("curved forward projection with the element's y sign flipped",
     "dy = -sdd * np.sin(gp.astype(np.float64)).astype(dtype)",
     "dy = sdd * np.sin(gp.astype(np.float64)).astype(dtype)",
     [P.pin_curved_fp_blob_on_arc]),
    ("cone forward projection without the step length",
     "val = ((T(1) - zw) * lo + zw * hi) * step * diag", "val = ((T(1) - zw) * lo + zw * hi) * step",
     [P.pin_cone_fp_is_trilinear_line_sum, P.pin_cone_fp_of_ball_is_chord]),
    ("helical travel from n_v_full - 1 rows",
     "(np.abs(v[1] - v[0]) * T(n_v_full))", "(np.abs(v[1] - v[0]) * T(n_v_full - 1))",
     [P.pin_helical_travel_is_pitch_times_isocentre_coverage_per_lap]),
    ("helical travel without the sid / sdd demagnification",
     "/ (T(sdd) / T(sid)))", "* (T(sdd) / T(sid)))",
     [P.pin_helical_travel_is_pitch_times_isocentre_coverage_per_lap]),
    ("helical forward projection without the gantry travel",
     "zs = zpv * J.astype(dtype)", "zs = zpv * J.astype(dtype) * T(0)",
     [P.pin_helical_fp_view_equals_circular_fp_of_the_shifted_volume]),
    ("helical rays aimed from the unmoved source height",
     "ex, ey, ez = dx - sid, dy - T(0), dz - zs", "ex, ey, ez = dx - sid, dy - T(0), dz",
     [P.pin_helical_fp_view_equals_circular_fp_of_the_shifted_volume]),
    ("fan back-projection with the view weight squared",
     "yield ok, (sid * w) * val * db[j] * div", "yield ok, (sid * w) * val * db[j] * db[j] * div",
     [P.pin_fbp_of_fp_reproduces_phantom]),
]


def _mutant(old, new):
    assert SRC.count(old) == 1, f"planted text must occur once in oracle/recon.py: {old!r}"
    mod = types.ModuleType("recon_mutant")
    mod.__file__ = "recon_mutant.py"
    exec(compile(SRC.replace(old, new), "recon_mutant.py", "exec"), mod.__dict__)
    return mod


@pytest.mark.parametrize("name,old,new,pins", MUTATIONS, ids=[m[0] for m in MUTATIONS])
def test_pins_catch_planted_mistake(name, old, new, pins):
    mod = _mutant(old, new)
    for pin in pins:
        with pytest.raises(AssertionError):
            pin(mod)
