"""The RASPA2-side C binding (context creation, energies at a position,
Widom moves) as a compiled C program against include/rhg.h, linked with
librhg.so: it compiles with
-Wall -Werror as C99, refuses force fields whose pair table is not a
Lorentz-Berthelot mix, and (on the GPU) gives the energies and Widom block
sums of the Python binding of the same library."""
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "c_binding", "binding.c")
PREBUILT = os.path.join(HERE, "c_binding", "binding")
K2E = 0.8314464919  # the program's KELVIN_TO_ENERGY


def compile_binding(out):
    subprocess.check_call(["gcc", "-std=c99", "-O2", "-Wall", "-Wextra", "-Werror",
                           "-I" + os.path.join(ROOT, "include"), SRC,
                           "-L" + os.path.join(ROOT, "raspa2_amd"),
                           "-Wl,-rpath," + os.path.join(ROOT, "raspa2_amd"), "-lrhg", "-lm", "-o", out])
    return out


def deck(W, path, override=False, n_pos=64):
    """The 25.8 A charged cell with four framework pseudo-atom types
    (atom i has type i % 4) and the CO2 probe's two types; override=True
    puts an explicit pair line (not the mix) between framework type 1 and
    the probe's oxygen."""
    fw = W.c1_framework(n_rep=1)
    ftypes = [(40.0, 3.0, 0.3), (60.0, 3.2, -0.3), (25.0, 2.7, 0.15), (70.0, 3.4, -0.15)]
    ptypes = [(79.0, 3.05, -0.35), (27.0, 2.80, 0.70)]
    types = [t + (1,) for t in ftypes] + [t + (0,) for t in ptypes]
    T = len(types)
    lines = [f"{fw.n} {T} 1", " ".join(repr(float(v)) for v in np.asarray(fw.box).reshape(9)),
             "12.0 300.0 1"]
    lines += [f"{e!r} {s!r} {q!r} {f}" for e, s, q, f in types]
    for i in range(T):
        for j in range(T):
            e = float(np.sqrt(types[i][0] * types[j][0]))
            s = 0.5 * (types[i][1] + types[j][1])
            if override and {i, j} == {1, 4}:
                e *= 1.2
            lines.append(f"1 {e!r} {s!r}")
    for i, p in enumerate(fw.xyz):
        lines.append(f"{float(p[0])!r} {float(p[1])!r} {float(p[2])!r} {i % 4}")
    lines.append("3 1 1 10")
    for t, z in ((4, 1.16), (5, 0.0), (4, -1.16)):
        lines.append(f"{t} 0.0 0.0 {z!r}")
    rng = np.random.default_rng(3)
    pts = rng.uniform(0, 1, (n_pos, 3)) @ np.asarray(fw.box).reshape(3, 3)
    lines.append(str(n_pos))
    lines += [f"{float(p[0])!r} {float(p[1])!r} {float(p[2])!r}" for p in pts]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return fw, ftypes, ptypes, pts


def test_binding_compiles_links_and_refuses_overrides(W, R, tmp_path):
    exe = compile_binding(str(tmp_path / "binding"))
    for override, word in ((False, "ok"), (True, "refused")):
        d = str(tmp_path / f"deck{int(override)}.txt")
        deck(W, d, override=override)
        out = subprocess.run([exe, d, "--host-only"], capture_output=True, text=True, check=True).stdout
        assert "gfx950" in out and f"mixing {word}" in out
        rc, w = out.split("estimators ")[1].split()
        b = R.WidomBlocks()
        for k in range(5):
            b.count[k] = b.ideal_gas[k] = 100.0
            b.sum_w[k] = 1.0 + k
        assert int(rc) == 0 and float(w) == R.widom_estimators(b, 300.0, 1000.0)["w_overall"]


@pytest.mark.gpu
def test_binding_matches_python_path(W, R, tmp_path):
    exe = PREBUILT if os.path.exists(PREBUILT) else compile_binding(str(tmp_path / "binding"))
    d = str(tmp_path / "deck.txt")
    fw, ftypes, ptypes, pts = deck(W, d)
    print("C binding: 576-atom charged cell, CO2", flush=True)
    out = subprocess.run([exe, d], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.strip().splitlines()
    e_c = np.array([[float(v) for v in line.split()] for line in lines[:len(pts)]])
    blocks_c = np.array([[float(v) for v in line.split()] for line in lines[len(pts):]])
    # the same system through the Python binding (eps through the program's
    # K -> internal -> K round trip)
    ty = np.arange(fw.n) % 4
    eps = np.array([ftypes[t][0] * K2E / K2E for t in ty])
    ctx = R.Context(fw.box, fw.bc, fw.xyz[:, 0], fw.xyz[:, 1], fw.xyz[:, 2], eps,
                    np.array([ftypes[t][1] for t in ty]), charge=np.array([ftypes[t][2] for t in ty]),
                    charge_method=R.CHARGE_EWALD, precision=R.PREC_FP64)
    o = ctx.add_pseudo_atom(ptypes[0][0] * K2E / K2E, ptypes[0][1], ptypes[0][2])
    c = ctx.add_pseudo_atom(ptypes[1][0] * K2E / K2E, ptypes[1][1], ptypes[1][2])
    uv, uc, _ = ctx.energy_at_positions(pts, c)
    ok = uv < 1e7
    assert np.allclose(e_c[ok, 0], uv[ok], rtol=1e-12, atol=1e-9)
    assert np.allclose(e_c[:, 1], uc, rtol=1e-12, atol=1e-9)
    mol = R.molecule([o, c, o], np.array([[0, 0, 1.16], [0, 0, 0], [0, 0, -1.16]]), start=1, k_fb=1, k_or=10)
    b, _ = ctx.widom_run(mol, 4242, 20000, steps_per_cycle=20)
    assert np.allclose(blocks_c, b.as_array(), rtol=1e-10, atol=0)
    ctx.close()
