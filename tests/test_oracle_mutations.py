"""Mutation check of the oracle's known-answer tests.

Each case plants one plausible mistake in a copy of oracle/oracle.cpp (a swapped
operand, a dropped factor, a wrong comparison or random-number slot), builds that
copy on the CPU and runs tests/test_oracle_kat.py against it in a child process
(PRT_ORACLE_LIB points the oracle's loader at the copy). The mistake must make at
least one named test fail: otherwise the tests would not pin that part of the
estimator (src/renderer/renderer.cu:158-161, :171-205, :239-243, :253-259,
:375-393, :464-493).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
SRC = ROOT / "oracle" / "oracle.cpp"

# name -> (original text, planted text, tests that must fail)
MUTATIONS = {
    "refraction lobe test against P, not 0.25": (
        "if (T(smp.draw(blockA(b), 0)) > T(0.25)) {", "if (T(smp.draw(blockA(b), 0)) > P) {",
        ["test_refractive_normal_incidence_takes_one_of_two_closed_forms"]),
    "RP and TP swapped": (
        "mask = mask * load3<T>(mat.ks) * TP;\n                } else {\n                    mask = mask * RP;",
        "mask = mask * load3<T>(mat.ks) * RP;\n                } else {\n                    mask = mask * TP;",
        ["test_refractive_normal_incidence_takes_one_of_two_closed_forms"]),
    "ks dropped on transmission": (
        "mask = mask * load3<T>(mat.ks) * TP;", "mask = mask * TP;",
        ["test_refractive_normal_incidence_takes_one_of_two_closed_forms"]),
    "Fresnel R0 with the precedence corrected": (
        "T R0 = (etai - etat) * (etai - etat) / (etat + etai) * (etat + etai);",
        "T R0 = (etai - etat) * (etai - etat) / ((etat + etai) * (etat + etai));",
        ["test_fresnel_quirk_normal_incidence", "test_refractive_normal_incidence_takes_one_of_two_closed_forms"]),
    "no emission after a refractive vertex": (
        "if (b == 0 || previousBsdf == PRT_SPECULAR || previousBsdf == PRT_REFRACTIVE) {",
        "if (b == 0 || previousBsdf == PRT_SPECULAR) {",
        ["test_refractive_normal_incidence_takes_one_of_two_closed_forms"]),
    "roulette survivors not divided by p": (
        "            mask = mask / p;\n        }\n        if (kAdjoint && b >= cfg.rr_start)",
        "        }\n        if (kAdjoint && b >= cfg.rr_start)",
        ["test_roulette_kills_by_its_own_draw_and_rescales_survivors"]),
    "roulette draw from the spare slot": (
        "if (T(smp.draw(blockB(b), 2)) > p) break;", "if (T(smp.draw(blockB(b), 3)) > p) break;",
        ["test_roulette_kills_by_its_own_draw_and_rescales_survivors",
         "test_roulette_starts_at_bounce_three_between_two_mirrors"]),
    "roulette draw from the BSDF block": (
        "if (T(smp.draw(blockB(b), 2)) > p) break;", "if (T(smp.draw(blockA(b), 2)) > p) break;",
        ["test_roulette_kills_by_its_own_draw_and_rescales_survivors"]),
    "roulette from bounce 4": (
        "if (b >= cfg.rr_start) {\n            p = maxComponent", "if (b > cfg.rr_start) {\n            p = maxComponent",
        ["test_roulette_starts_at_bounce_three_between_two_mirrors"]),
    "DIFFSPEC diffuse weight 1/p dropped": (
        "currentBsdf = PRT_DIFFUSE;\n                mask = mask * (T(1) / p);", "currentBsdf = PRT_DIFFUSE;",
        ["test_diffspec_is_the_weighted_mixture_of_its_lobes"]),
    "DIFFSPEC specular weight 1/(1-p) dropped": (
        "mask = mask * (T(1) / (T(1) - p));", "",
        ["test_diffspec_is_the_weighted_mixture_of_its_lobes"]),
    "DIFFSPEC lobe comparison reversed": (
        "if (T(smp.draw(blockA(b), 0)) < p) {", "if (T(smp.draw(blockA(b), 0)) > p) {",
        ["test_diffspec_is_the_weighted_mixture_of_its_lobes"]),
    "specular ks dropped": (
        "mask = mask * load3<T>(mat.ks) / it.pdf;", "mask = mask / it.pdf;",
        ["test_mirror_then_emitter_is_ks_times_3ka"]),
    "mirror about the negated normal": (
        "it.inputDirection = o - n * dot(o, n) * T(2);", "it.inputDirection = o + n * dot(o, n) * T(2);",
        ["test_mirror_then_emitter_is_ks_times_3ka"]),
    "diffuse pdf 1 instead of 0.5": (
        "            it.pdf = T(0.5);\n            T cw", "            it.pdf = T(1);\n            T cw",
        ["test_next_event_estimate_expectation"]),
    "diffuse cosine sqrt(r2) for sqrt(1 - r2)": (
        "w * std::sqrt(T(1) - r2));\n}", "w * std::sqrt(r2));\n}",
        ["test_diffuse_sampler_moments"]),
}


def _run_mutant(name: str | None, tmp: Path) -> tuple[str | None, list[str], str]:
    """Builds the oracle with mutation `name` (None: unchanged) and runs the tests."""
    src = SRC.read_text()
    if name is not None:
        old, new, _ = MUTATIONS[name]
        assert src.count(old) == 1, f"{name}: the original text is not unique in oracle.cpp"
        src = src.replace(old, new)
    d = tmp / f"m{abs(hash(name)) % 10**8}"
    d.mkdir()
    (d / "oracle.cpp").write_text(src)
    lib = d / "liboracle.so"
    subprocess.run(["g++", "-O1", "-fPIC", "-fopenmp", "-std=c++17", "-ffp-contract=off", f"-I{ROOT / 'include'}",
                    f"-I{ROOT / 'oracle'}", "-shared", str(d / "oracle.cpp"), "-o", str(lib)], check=True)
    env = dict(os.environ, PRT_ORACLE_LIB=str(lib), OMP_NUM_THREADS="2")
    p = subprocess.run([sys.executable, "-m", "pytest", str(ROOT / "tests" / "test_oracle_kat.py"), "-q",
                        "-p", "no:cacheprovider"], cwd=ROOT, env=env,
                       capture_output=True, text=True, timeout=600)
    failed = [l.split("::")[1].split(" ")[0] for l in p.stdout.splitlines() if l.startswith("FAILED")]
    return name, failed, p.stdout[-2000:]


def test_every_planted_mistake_fails_a_known_answer_test(tmp_path):
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 2)) as ex:
        results = list(ex.map(lambda n: _run_mutant(n, tmp_path), [None, *MUTATIONS]))
    _, failed, out = results[0]
    assert failed == [] and " passed" in out, f"the unchanged copy must pass every test:\n{out}"
    missed = []
    for name, failed, out in results[1:]:
        want = MUTATIONS[name][2]
        if not set(want) <= set(failed):
            missed.append(f"{name}: expected {want} to fail, failed {failed}\n{out}")
    assert not missed, "\n".join(missed)
