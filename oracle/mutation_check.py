#!/usr/bin/env python3
"""Mutation check of the oracle's pins (test infrastructure).

Each entry below is a plausible mistake in a restatement of the reference: a wrong
normalisation, a swapped operand, a wrong stencil origin or sign.  For each, this script builds
a copy of pond_oracle.c with that one change into a temporary directory and runs
tests/test_oracle_pins.py against it (pyoracle loads the build named by POND_ORACLE_LIB).  A
mutation the pins do not catch is reported as MISSED and the script exits non-zero.

    python oracle/mutation_check.py [NAME ...]
"""
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

# name -> (text in pond_oracle.c, replacement); the reference line each one breaks in brackets
MUTATIONS = {
    # WENO5 smoothness indicator beta1 normalised by 6 instead of 3 (steps.cu:804)
    "beta1_over_6": ("+ 13 * u3 * u3 - 13 * u2 * u3 + 4 * u2 * u2) / 3.0;", "+ 13 * u3 * u3 - 13 * u2 * u3 + 4 * u2 * u2) / 6.0;"),
    # beta2 used in place of beta0 (steps.cu:809-811)
    "beta0_beta2_swap": ("w0 /= (sc + beta0) * (sc + beta0);", "w0 /= (sc + beta2) * (sc + beta2);"),
    # linear weights d0 and d2 exchanged (steps.cu:796-798)
    "d0_d2_swap": ("const double d0 = (cdist * cdist - 3 * cdist + 2) / 12.;",
                   "const double d0 = (cdist * cdist + 3 * cdist + 2) / 12.;"),
    # as-written y/z sub-stencil moved from floor(y+0.5)+0..2 to -1..+1 (steps.cu:776-778)
    "weno5_yz_centred": ("const int ly = centred ? ry - 1 : ry, lz = centred ? rz - 1 : rz;",
                         "const int ly = ry - 1, lz = rz - 1;"),
    # smoothness row through the departure point instead of (r mod Ny, r mod Nz) (steps.cu:783)
    "weno_row_quirk_off": ("const long by = centred ? wrap(ry, S->p.ny) : wrap(r, S->p.ny);",
                           "const long by = wrap(ry, S->p.ny);"),
    # generic WENO5: indicator of the middle stencil normalised by 6 (steps.cu:643)
    "weno5g_beta1_over_6": ("+ 5 * u2 * u4 - 13 * u3 * u4 + 4 * u4 * u4) / 3.0;", "+ 5 * u2 * u4 - 13 * u3 * u4 + 4 * u4 * u4) / 6.0;"),
    # generic WENO5: linear weight of the left stencil on the right one (steps.cu:621)
    "weno5g_d0_sign": ("const double d0 = (cdist * cdist + 3 * cdist + 2) / 12.;",
                       "const double d0 = (cdist * cdist - 3 * cdist + 2) / 12.;"),
    # generic WENO5: lshift = 0, the commented-out choice (steps.cu:614-615)
    "weno5g_lshift0": ("const int lshift = 1;", "const int lshift = 0;"),
    # generic WENO5: weight of S0 from the left stencil's indicator (steps.cu:645)
    "weno5g_w0_beta2": ("double w0 = d0 / ((1e-6 + beta0) * (1e-6 + beta0));",
                        "double w0 = d0 / ((1e-6 + beta2) * (1e-6 + beta2));"),
    # AMR_remove: neighbours up to 4 cells away instead of 3 (amr.cu:50)
    "amr_remove_bound_4": ("(distM < 4 && distP < 4)", "(distM <= 4 && distP < 4)"),
    # AMR_add: new cell one past the middle of the gap (amr.cu:15)
    "amr_add_mid_offset": ("(x + dist / 2) % nx", "(x + dist / 2 + 1) % nx"),
    # AMR_add: criterion inverted (amr.cu:16)
    "amr_add_criterion": ("> 0.01 || fabs(1 - rec_rho", "< 0.01 || fabs(1 - rec_rho"),
    # AMR_add: the new cell's density from one side only (amr.cu:19)
    "amr_add_one_sided": ("c0.rho = 0.5 * (cm.rho + cp.rho);", "c0.rho = 0.5 * (cm.rho + cm.rho);"),
    # LBMpull steps switched-off cells (steps.cu:359)
    "amr_step_ignores_set": ("held (LBMpull, steps.cu:359) */\n  if (!S->set[idx]) return;",
                             "held (LBMpull, steps.cu:359) */\n  if (0) return;"),
    # D3Q125 interpolation origin half a cell off in y (steps.cu:163,209)
    "q125_origin_half_y": ("crd[1] - (double)(y - 1), crd[2] - (double)(z - 1)};",
                           "crd[1] - (double)(y - 1) + 0.5, crd[2] - (double)(z - 1)};"),
    # D3Q125 departure point at x + v (steps.cu:207-208)
    "q125_x_plus_v": ("const double crd[3] = {(double)x - vi[inew][0], (double)y - vi[inew][1], (double)z - vi[inew][2]};",
                      "const double crd[3] = {(double)x + vi[inew][0], (double)y + vi[inew][1], (double)z + vi[inew][2]};"),
    # D3Q27 / D1Q5 departure point at x + v along x (steps.cu:444)
    "q27_x_plus_v": ("const double crd[3] = {(double)x - vi[i][0], (double)y - vi[i][1], (double)z - vi[i][2]};",
                     "const double crd[3] = {(double)x + vi[i][0], (double)y - vi[i][1], (double)z - vi[i][2]};"),
    # energy total with 2 T instead of Dim T (update.cu:67)
    "totals_energy_dim2": ("en += rho * (ux * ux + uy * uy + uz * uz + Dim * T) * 1.0;",
                           "en += rho * (ux * ux + uy * uy + uz * uz + 2 * T) * 1.0;"),
    # D3Q27 transfer: resting-axis basis A^2 + T1 instead of A^2 - T1 (structs.cuh:536-538)
    "q27_rest_basis_sign": ("const double gx = Ax * Ax - Bx;", "const double gx = Ax * Ax + Bx;"),
    # mirror boundary: low face reflected to -p instead of -1-p (steps.cu:1008)
    "mirror_reflect_low": ("if (p < 0) { p = -1 - p; *rev = 1; }", "if (p < 0) { p = -p; *rev = 1; }"),
    # mirror boundary: high face reflected to 2N-p instead of 2N-1-p (steps.cu:1009)
    "mirror_reflect_high": ("if (p >= n) { p = n - 1 - p + n; *rev = 1; }", "if (p >= n) { p = n - p + n; *rev = 1; }"),
    # Cell::mirrX without the velocity sign flip (structs.cuh:553)
    "mirror_no_u_flip": ("  uT[d] = -uT[d];\n", "  uT[d] = uT[d];\n"),
    # Cell::mirrX without the population exchange (structs.cuh:555-557)
    "mirror_no_permutation": ("fn[i] = L->ef[i][d] == 0 ? f[i] : f[L->rev[d][i]];", "fn[i] = f[i];"),
    # calcMoments: temperature without the |u|^2 subtraction of one component (structs.cuh:330)
    "moments_drop_uz2": ("u3 = (u3 - u0 * u0 - u1 * u1 - u2 * u2) * (1. / L->Dim);",
                         "u3 = (u3 - u0 * u0 - u1 * u1) * (1. / L->Dim);"),
}


def main():
    src = open(os.path.join(HERE, "pond_oracle.c")).read()
    missed = []
    only = sys.argv[1:]  # optional subset of mutation names
    with tempfile.TemporaryDirectory() as tmp:
        for name, (a, b) in MUTATIONS.items():
            if only and name not in only:
                continue
            if src.count(a) != 1:
                print(f"{name}: pattern not found exactly once in pond_oracle.c", flush=True)
                missed.append(name)
                continue
            csrc = os.path.join(tmp, f"{name}.c")
            lib = os.path.join(tmp, f"{name}.so")
            with open(csrc, "w") as fh:
                fh.write(src.replace(a, b))
            subprocess.run(["gcc", "-std=gnu11", "-O3", "-fPIC", "-fopenmp", "-ffp-contract=off", "-fno-fast-math",
                            "-shared", "-o", lib, csrc, os.path.join(HERE, "mlp_oracle.c"), "-lm"], check=True)
            env = dict(os.environ, POND_ORACLE_LIB=lib)
            r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                                os.path.join(ROOT, "tests", "test_oracle_pins.py")],
                               capture_output=True, text=True, cwd=ROOT, env=env)
            failed = [ln.split(" ")[1] for ln in r.stdout.splitlines() if ln.startswith("FAILED")]
            caught = r.returncode != 0 and failed
            print(f"{name}: {'CAUGHT by ' + failed[0] if caught else 'MISSED'}", flush=True)
            if not caught:
                missed.append(name)
    total = len(only) if only else len(MUTATIONS)
    print(f"{total - len(missed)}/{total} mutations caught")
    sys.exit(1 if missed else 0)


if __name__ == "__main__":
    main()
