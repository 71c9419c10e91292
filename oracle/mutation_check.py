"""Mutation check of the oracle's pins (test infrastructure, CPU only).

Plants one mistake at a time in a copy of oracle/refsim.cpp (a dropped
term, a flipped sign, an i-for-j index, a transposed operand, a dropped
factor), builds the copy in a temporary directory, and runs the pin tests
(tests/test_oracle_pins.py, tests/test_oracle.py, tests/test_segment.py) against it through
DIMCV_ORACLE_LIB.  Every mutation must make a pin fail; the script prints
which test caught each and exits 1 if any survived.  The sources under
oracle/ are never modified.

    python oracle/mutation_check.py [-j 4]
"""
import argparse
import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

# (name, text in refsim.cpp (must occur once), replacement)
MUTATIONS = [
 ("density self term", "    float dens = s->m_V[p_i] * s->W(0.0f);\n    float den = 0.0f;", "    float dens = 0.0f;\n    float den = 0.0f;"),
 ("density solids", "      else if (s->is_solid(p_j)) den += s->m_V[p_j] * s->W(vnorm(s->rel(xi, s->X(p_j))));", "      else if (s->is_solid(p_j)) den += 0.0f;"),
 ("factor sum of squares", "        r3 += vdot(g, g);", "        r3 += 0.0f;"),
 ("factor square of sum", "    float sum = r3;\n    V3 gp{r0, r1, r2};\n    sum += vdot(gp, gp);", "    float sum = r3;\n    V3 gp{r0, r1, r2};\n    sum += 0.0f * vdot(gp, gp);"),
 ("factor solid weight", "        V3 g = smul(s->m_V[p_j] * s->f_density0, s->gradW(s->rel(xi, s->X(p_j))));\n        r0 += g.x;", "        V3 g = smul(s->m_V[p_j], s->gradW(s->rel(xi, s->X(p_j))));\n        r0 += g.x;"),
 ("rho-dot sign", "      if (s->is_fluid(p_j))\n        dadv += s->m_V[p_j] * vdot(vsub(vi, vj),", "      if (s->is_fluid(p_j))\n        dadv += s->m_V[p_j] * vdot(vsub(vj, vi),"),
 ("rho-dot 20 rule", "    if (nn < 20) d = 0.0f;", "    if (nn < 2) d = 0.0f;"),
 ("div update sign", "          dv = vsub(dv, smul(dt * k_sum, g));", "          dv = vadd(dv, smul(dt * k_sum, g));"),
 ("div solid sign", "          V3 vc = smul(-dt * 1.0f * k_i, g);\n          dv = vadd(dv, vc);", "          V3 vc = smul(dt * 1.0f * k_i, g);\n          dv = vadd(dv, vc);"),
 ("div rho_j -> rho_i", "float k_j = b_j * s->dfsph_factor[p_j] * s->density[p_j];", "float k_j = b_j * s->dfsph_factor[p_j] * s->density[p_i];"),
 ("div m_j -> m_i", "V3 g = smul(s->m[p_j], s->gradW(s->rel(xi, s->X(p_j))));\n          dv =", "V3 g = smul(s->m[p_i], s->gradW(s->rel(xi, s->X(p_j))));\n          dv ="),
 ("transit dropped", "else vi = vadd(vi, smul(s->transit_weight(xi.x), dv));", "else vi = vadd(vi, dv);"),
 ("visc sign", "V3 fv = smul(s->f_visc_coef * s->m_V[p_j]", "V3 fv = smul(-s->f_visc_coef * s->m_V[p_j]"),
 ("visc h2 dropped", "(rn * rn + s->f_visc_h2)", "(rn * rn + 1e-12f)"),
 ("vorticity operand order", "vort = vadd(vort, smul(s->m_V[p_j], vcross(vsub(s->Vv(p_j), vi),", "vort = vadd(vort, smul(s->m_V[p_j], vcross(vsub(vi, s->Vv(p_j)),"),
 ("grad v transposed", "G[3 * r + c] = G[3 * r + c] + av[r] * kv[c];", "G[3 * r + c] = G[3 * r + c] + av[c] * kv[r];"),
 ("d_vort sign", "Sim::put(s->d_vort, p_i, vsub(Sim::get(s->vorticity, p_i), curl));", "Sim::put(s->d_vort, p_i, vsub(curl, Sim::get(s->vorticity, p_i)));"),
 ("smoothing self term", "    V3 acc = vmuls(smul(s->m_V[p_i], Sim::get(s->d_vort, p_i)), w0);", "    V3 acc{0.0f, 0.0f, 0.0f};"),
 ("smoothing j -> i", "acc = vadd(acc, vmuls(smul(s->m_V[p_j], Sim::get(s->d_vort, p_j)),", "acc = vadd(acc, vmuls(smul(s->m_V[p_j], Sim::get(s->d_vort, p_i)),"),
 ("KVN inverted", "float kvn = std::sqrt(Wn) / (std::sqrt(Sn) + 1e-8f);", "float kvn = std::sqrt(Sn) / (std::sqrt(Wn) + 1e-8f);"),
 ("production complement", "float produce = 1.0f - 1.0f / powf_c(kvn + 1.0f, rr);", "float produce = 1.0f / powf_c(kvn + 1.0f, rr);"),
 ("deletion complement", "float del = expf_c(s->f_neg_b * kvn);", "float del = 1.0f - expf_c(s->f_neg_b * kvn);"),
 ("AKDE self term", "    float kden = s->kvn[p_i] * e0;", "    float kden = 0.0f;"),
 ("Epanechnikov square", "if (q[d] < 1.0f) ret *= 0.75f * (1.0f - q[d] * q[d]);", "if (q[d] < 1.0f) ret *= 0.75f * (1.0f - q[d]);"),
 ("rho* dt dropped", "    float da = s->density[p_i] + dt * delta;", "    float da = s->density[p_i] + delta;"),
 ("rho* solid weight", "        delta += s->m_V[p_j] * s->f_density0 * vdot(vsub(vi, vj),", "        delta += s->m_V[p_j] * vdot(vsub(vi, vj),"),
 ("press k_j index", "float k_j = b_j * s->dfsph_factor[p_j];", "float k_j = b_j * s->dfsph_factor[p_i];"),
 ("press update sign", "          if (!s->cfg.has_transit) vi = vsub(vi, smul(dt * k_sum, g));", "          if (!s->cfg.has_transit) vi = vadd(vi, smul(dt * k_sum, g));"),
 ("vis normaliser", "Sim::put(s->vorticity_vis, p_i, vdivs(vort, w + 1e-6f));", "Sim::put(s->vorticity_vis, p_i, vdivs(vort, 1.0f + 1e-6f));"),
 ("vis self weight", "    V3 vort = smul(w, Sim::get(s->vorticity, p_i));", "    V3 vort = smul(w, Sim::get(s->vorticity, p_i));\n    w = 0.0f;"),
 ("boundary volume factor", "    s->m_V[p_i] = 1.0f / delta * 3.0f;\n  }\n}\n\n// sph_base.py:105-112", "    s->m_V[p_i] = 1.0f / delta * 2.0f;\n  }\n}\n\n// sph_base.py:105-112"),
 ("bounce restitution", "      const float c_f = 0.5f;", "      const float c_f = 0.0f;"),
 ("bounce clamp side", "      if (q[d] <= pad) {\n        n[d] += -1.0f;", "      if (q[d] <= pad) {\n        n[d] += 1.0f;"),
 ("near cross order", "      vv = vadd(vv, smul(A, vcross(r, Sim::get(s->d_vort_smoothed, p_j))));", "      vv = vadd(vv, smul(A, vcross(Sim::get(s->d_vort_smoothed, p_j), r)));"),
 ("near 3-D norm", "      V3 r{r2.x, r2.y, 0.0f};\n      float rn = vnorm(r);", "      V3 r{r2.x, r2.y, 0.0f};\n      float rn = vnorm(r2);"),
 ("far kden divisor dropped", "      t = vdivs(t, s->kernel_density[p_j] + 1e-8f);", "      t = vdivs(t, 1.0f + 1e-8f);"),
 ("far count", "else vs = vmuls(vs, fac / (float)added);", "else vs = vmuls(vs, fac / (float)(added + 1));"),
 ("far S_i", "  float S = (float)(s->kden_sum - (double)kns);", "  float S = (float)(s->kden_sum);"),
 ("rigid reaction sign", "V3 add{((-vc.x) * s->f_inv_dt) * ri / rj,", "V3 add{((vc.x) * s->f_inv_dt) * ri / rj,"),
 ("segment cross order", "um = vadd(um, vdivs(smul(c->inv4pi, vcross(w, r)), den));", "um = vadd(um, vdivs(smul(c->inv4pi, vcross(r, w)), den));"),
 ("segment regularisation", "      float r2 = vdot(r, r) + R2;", "      float r2 = vdot(r, r);"),
 ("ring core dropped", "double core = 1.0 - std::exp(-r2 * is2);", "double core = 1.0;"),
]


def run_one(mut, tmp_root):
    name, a, b = mut
    src = open(os.path.join(HERE, "refsim.cpp")).read()
    if src.count(a) != 1:
        return name, f"pattern occurs {src.count(a)} times"
    d = tempfile.mkdtemp(dir=tmp_root)
    for f in ("frozen.cpp", "refmath.h", "Makefile"):
        shutil.copy(os.path.join(HERE, f), d)
    with open(os.path.join(d, "refsim.cpp"), "w") as fh:
        fh.write(src.replace(a, b, 1))
    subprocess.run(["make", "-s", "-C", d], check=True, capture_output=True)
    env = dict(os.environ, DIMCV_ORACLE_LIB=os.path.join(d, "liboracle.so"), OMP_NUM_THREADS="2")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        "tests/test_oracle_pins.py", "tests/test_oracle.py", "tests/test_segment.py"],
                       cwd=ROOT, env=env, capture_output=True, text=True)
    failed = [ln for ln in r.stdout.splitlines() if ln.startswith("FAILED")]
    if r.returncode == 0:
        return name, "SURVIVED"
    return name, "caught by " + (failed[0].split("::")[1].split(" ")[0] if failed else "an error")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=4)
    args = ap.parse_args()
    tmp_root = tempfile.mkdtemp(prefix="oracle_mut_")
    try:
        with ThreadPoolExecutor(args.j) as ex:
            results = list(ex.map(lambda m: run_one(m, tmp_root), MUTATIONS))
    finally:
        shutil.rmtree(tmp_root, ignore_errors=True)
    bad = 0
    for name, outcome in results:
        print(f"{name:28s} {outcome}")
        bad += not outcome.startswith("caught")
    print(f"{len(results) - bad}/{len(results)} mutations caught")
    return 1 if bad else 0


if __name__ == "__main__":
    sys.exit(main())
