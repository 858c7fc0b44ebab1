"""The oracle's pins catch the mistakes they are meant to catch.

Each case copies the oracle's sources, plants one plausible misreading of the
reference into the copy (a swapped bound, a transposed operand, a dropped
threshold branch, a swapped blend), builds it with the oracle's own flags and
runs the named pin tests against it in a child process
(MPM_ORACLE_LIB points pyoracle at the copy).  Every pin must fail on its
mutant; the unmodified oracle passes the same tests (tests/test_oracle_pins.py).
CPU only: the mistakes live in the oracle, never in GPU code.
"""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-shared"]

# (name, text in mpm_oracle.hpp, replacement, pin tests that must fail)
MUTANTS = [
    ("clamp_bounds_swapped",
     "R hi = R(1.0f) + p.crit_s, lo = R(1.0f) - p.crit_c;",
     "R hi = R(1.0f) + p.crit_c, lo = R(1.0f) - p.crit_s;",
     "test_update_F_clamps_to_the_named_bound"),
    ("Fp_transposed_operand",
     "p.Fp = mm(mm(mm(V, inv(S)), tr(U)), F);",
     "p.Fp = mm(mm(mm(U, inv(S)), tr(V)), F);",
     "test_update_F_product_invariant_and_clamp_range"),
    ("Fe_trial_wrong_side",
     "p.Fe = mm(madd(eye<R>(), mscale(dt, grad_v)), p.Fe);",
     "p.Fe = mm(p.Fe, madd(eye<R>(), mscale(dt, grad_v)));",
     "test_update_F_product_invariant_and_clamp_range"),
    ("blend_swapped",
     "else vel = add(scale(blend_, vflip), scale(R(1.0 - blend_), vpic));",
     "else vel = add(scale(blend_, vpic), scale(R(1.0 - blend_), vflip));",
     "test_pic_flip_blend_two_lattices"),
    ("flip_increment_sign",
     "vflip = add(vflip, scale(w, sub(n.velocity_star, n.velocity)));",
     "vflip = add(vflip, scale(w, sub(n.velocity, n.velocity_star)));",
     "test_pic_flip_blend_two_lattices"),
    ("light_nodes_divided",
     "            if (n.mass > R(kFltEps)) {\n                for (int i = 0; i < 3; ++i) n.velocity[i] = n.velocity[i] / n.mass;",
     "            if (n.mass > R(0)) {\n                for (int i = 0; i < 3; ++i) n.velocity[i] = n.velocity[i] / n.mass;",
     "test_light_node_keeps_momentum_no_gravity or test_light_node_branch_in_one_step"),
    ("light_nodes_get_gravity",
     "if (n.mass > R(kFltEps)) n.force[1] = R(double(n.force[1]) - double(gravity) * double(n.mass));",
     "n.force[1] = R(double(n.force[1]) - double(gravity) * double(n.mass));",
     "test_light_node_keeps_momentum_no_gravity"),
]


def _build(tmp, name, old=None, new=None):
    """Copy of the oracle's sources (with one text replaced, if given), built
    as a separate liboracle_mpm.so; returns its path."""
    src = open(os.path.join(ORACLE, "mpm_oracle.hpp")).read()
    if old is not None:
        assert src.count(old) == 1, f"{name}: mutation site not found exactly once"
        src = src.replace(old, new)
    d = os.path.join(tmp, name)
    os.makedirs(os.path.join(d, "oracle"))
    os.makedirs(os.path.join(d, "include"))
    shutil.copy(os.path.join(ROOT, "include", "mpm_hip.h"), os.path.join(d, "include"))
    shutil.copy(os.path.join(ORACLE, "oracle_capi.cpp"), os.path.join(d, "oracle"))
    with open(os.path.join(d, "oracle", "mpm_oracle.hpp"), "w") as f:
        f.write(src)
    lib = os.path.join(d, "liboracle_mpm.so")
    subprocess.run(["g++"] + FLAGS + ["-o", lib, os.path.join(d, "oracle", "oracle_capi.cpp")], check=True)
    return lib


def _run_pins(lib, select):
    env = dict(os.environ, MPM_ORACLE_LIB=lib)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider",
                        os.path.join(ROOT, "tests", "test_oracle_pins.py"), "-k", select],
                       cwd=ROOT, env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    return r.returncode, r.stdout


@pytest.mark.parametrize("name,old,new,select", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_pin_fails_on_mutant(tmp_path, name, old, new, select):
    lib = _build(str(tmp_path), name, old, new)
    rc, out = _run_pins(lib, select)
    tail = out.strip().splitlines()[-1] if out.strip() else ""
    print(f"[{name}] pins '{select}': rc={rc} {tail}", flush=True)
    assert rc == 1 and "failed" in tail, f"{name}: the pins did not catch the mutant\n{out[-2000:]}"


def test_pins_pass_on_unmodified_copy(tmp_path):
    """The same build recipe on unmodified sources passes every pin above,
    so a failure above is the mutation's doing."""
    lib = _build(str(tmp_path), "unmodified")
    select = " or ".join(sorted({s for m in MUTANTS for s in m[3].split(" or ")}))
    rc, out = _run_pins(lib, select)
    assert rc == 0, out[-2000:]
