"""CPU tests: planted slips in the SRRF oracle are caught by the closed forms.

Each case copies oracle/srrf_oracle.cpp, applies one plausible transcription
mistake (a dropped square or divisor, a wrong renormalisation, a flipped
sign, the wrong intensity factor, a reset or dropped ITC accumulator, ...),
compiles the copy into a temporary library and runs the closed-form check
that should catch it against that library: the check must fail.  The text
each slip replaces must occur exactly once, so that a change to the oracle
cannot make a case pass vacuously.
"""
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

import srrf_closed_forms as cf
from oracle import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "srrf_oracle.cpp")

# name -> (text in the oracle, replacement, checks that must fail)
RAD = [lambda: cf.check_radiality("d26"), lambda: cf.check_radiality("ring2")]
MUTANTS = {
    "drop_dk_square": ("                Dk = Dk * Dk;\n", "", RAD),
    "drop_divisor": ("            DivD = DivD / j.divisor;\n", "", RAD),
    "renormalise_without_half": ("DivD = (R)0.5 + (DivD / (R)2);", "DivD = (R)0.5 + DivD;",
                                 [lambda: cf.check_radiality("d26", renorm=1)]),
    "drop_positivity": ("R CGH = (j.c.RemoveCon == 1) ? std::fmax(DivD, (R)0) : DivD;", "R CGH = DivD;",
                        [lambda: cf.check_radiality("d26", poscon=1)]),
    "swap_sign": ("if (diverging) DivD -= Dk;", "if (!diverging) DivD -= Dk;", RAD),
    "divide_by_gmag_once": ("/ (GMag * GMag);", "/ GMag;", RAD),
    "drop_radius_scale": ("calc_dk(x0, y0, z0, Xc, Yc, Zc, Gx, Gy, Gz, GMag) / radius;",
                          "calc_dk(x0, y0, z0, Xc, Yc, Zc, Gx, Gy, Gz, GMag);", RAD),
    "trm_without_intensity": ("MIP = std::fmax(iw == 1 ? rad(t) * I(t) : rad(t), MIP);",
                              "MIP = std::fmax(rad(t), MIP);", [lambda: cf.check_temporal_weighted(0)]),
    "tra_without_intensity": ("mean += (iw == 1) ? rad(t) * I(t) : rad(t);", "mean += rad(t);",
                              [lambda: cf.check_temporal_weighted(1)]),
    "intensity_not_clamped": ("return std::fmax(s, (R)0);", "return s;",
                              [lambda: cf.check_temporal_weighted(1)]),
    "trppm_weight_mean_rad_i": ("for (int t = 0; t < T; t++) IW += I(t) / (R)T;",
                                "for (int t = 0; t < T; t++) IW += rad(t) * I(t) / (R)T;",
                                [lambda: cf.check_temporal_weighted(2)]),
    "trppm_skipped_pairs_not_counted": ("                } else\n                    counter += T - t0;\n",
                                        "                }\n", [lambda: cf.check_temporal_weighted(2)]),
    "trac_weight_mean_i": ("                IW += rad(t) * I(t);\n", "                IW += I(t);\n",
                           [lambda: cf.check_temporal_weighted(3, 2)]),
    "trac_divide_by_terms": ("else acc = std::fabs(AB) / (R)T;", "else acc = std::fabs(AB) / (R)(T - ord);",
                             [lambda: cf.check_temporal_weighted(3, 2)]),
    "trac4_drop_term": ("acc = std::fabs(ABCD - AB * CD - AC * BD - AD * BC) / (R)T;",
                        "acc = std::fabs(ABCD - AB * CD - AC * BD) / (R)T;",
                        [lambda: cf.check_temporal_weighted(3, 4)]),
    "itc_reset_abc": ("                    AB = 0;\n", "                    AB = 0;\n                    ABC = 0;\n",
                      [lambda: cf.check_itc(3)]),
    "itc_reset_all": ("                    AB = 0;\n",
                      "                    AB = 0; ABCD = 0; CD = 0; AC = 0; BD = 0; AD = 0; BC = 0;\n",
                      [lambda: cf.check_itc(4)]),
    "itc_keep_ab": ("                    AB = 0;\n", "", [lambda: cf.check_itc(4)]),
    "itc4_drop_term": ("acc += std::fabs((ABCD - AB * CD - AC * BD - AD * BC)) / (R)nB;",
                       "acc += std::fabs((ABCD - AB * CD - AC * BD)) / (R)nB;", [lambda: cf.check_itc(4)]),
    "itc_bin_not_zeroed": ("                        ser[t] = 0;\n", "", [lambda: cf.check_itc(3), lambda: cf.check_itc(4)]),
}


def _build(tmpdir, name):
    old, new, _ = MUTANTS[name]
    text = open(SRC).read()
    assert text.count(old) == 1, "%s: the replaced text occurs %d times" % (name, text.count(old))
    src = os.path.join(tmpdir, "%s.cpp" % name)
    with open(src, "w") as f:
        f.write(text.replace(old, new))
    out = os.path.join(tmpdir, "lib%s.so" % name)
    r = subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-shared", "-pthread", "-ffp-contract=off", src, "-o",
                        out], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("srrf_mutants"))
    with ThreadPoolExecutor(max(1, min(8, os.cpu_count() or 1))) as ex:
        futs = {n: ex.submit(_build, d, n) for n in MUTANTS}
        return {n: f.result() for n, f in futs.items()}


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_planted_slip_is_caught(mutant_libs, name):
    checks = MUTANTS[name][2]
    with oracle.using(mutant_libs[name]):
        for check in checks:
            with pytest.raises(AssertionError):
                check()
    # the unmodified oracle passes the same checks
    for check in checks:
        check()
