"""The oracle's pins catch planted mistakes (CPU only).

Each mutation below is a one-token slip in a copy of oracle/dose_oracle.c of
the kind a transcription can make (a sign, an operand order, a constant, a
composition).  The copy is compiled into a temporary library and the pins of
tests/test_oracle_pins.py (plus the closed-form source and pencil checks of
tests/test_oracle.py) run against it; at least one must fail.  The unmodified
copy, compiled the same way, must pass them all.
"""
import os
import re
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

from oracle import oracle as O
from tests import test_oracle as T
from tests import test_oracle_pins as P

SRC = os.path.join(os.path.dirname(O.__file__), "dose_oracle.c")
CFLAGS = ["-O1", "-fopenmp", "-ffp-contract=off", "-fPIC", "-std=c11", "-shared"]

# (function, text in that function, replacement)
MUTATIONS = [
    ("orc_geom_init", "g->src[2] = -xg * g->sinta;", "g->src[2] = xg * g->sinta;"),
    ("img_to_head_f", "float sx = -g->sinta, cx = g->costa;", "float sx = g->sinta, cx = g->costa;"),
    ("img_to_head_f", "float xg = xt * cx - yt * sx;", "float xg = xt * cx + yt * sx;"),
    ("img_to_head_f", "sx = -g->sinca; cx = g->cosca;", "sx = g->sinca; cx = g->cosca;"),
    ("img_to_head_f", "h[0] = -xc;", "h[0] = xc;"),
    ("head_to_img_f", "float sx = g->sinca, cx = g->cosca;", "float sx = -g->sinca, cx = g->cosca;"),
    ("head_to_img_f", "float xz = -h[0], yz = h[2], zz = h[1];", "float xz = -h[0], yz = h[1], zz = h[2];"),
    ("img_to_head_d", "double sx = -(double)g->sinta, cx = g->costa;",
     "double sx = (double)g->sinta, cx = g->costa;"),
    ("img_to_head_d", "double xg = xt * cx - yt * sx,", "double xg = xt * cx + yt * sx,"),
    ("img_to_head_d", "sx = -(double)g->sinca; cx = g->cosca;", "sx = (double)g->sinca; cx = g->cosca;"),
    ("img_to_head_d", "h[0] = -xc; h[1] = zc; h[2] = yc;", "h[0] = -xc; h[1] = xc; h[2] = yc;"),
    ("head_to_img_d", "double xz = -h[0],", "double xz = h[0],"),
    ("head_to_img_d", "double sx = g->sinca, cx = g->cosca;", "double sx = -g->sinca, cx = g->cosca;"),
    ("head_to_img_d", "sx = g->sinta; cx = g->costa;", "sx = -g->sinta; cx = g->costa;"),
    ("kernel_tilt_d", "double ux = g->src[0] - p[0]", "double ux = p[0] - g->src[0]"),
    ("kernel_tilt_d", "double kx = ay * uz - az * uy", "double kx = az * uy - ay * uz"),
    ("kernel_tilt_d", "ky = az * ux - ax * uz", "ky = ax * uz - az * ux"),
    ("kernel_tilt_d", "double rx = ky * d[2] - kz * d[1]", "double rx = kz * d[1] - ky * d[2]"),
    ("kernel_tilt_d", "rx += d[0] * c;", "rx += d[0];"),
    ("kernel_tilt_d", "/ (1.0 + c);", "/ (1.0 - c);"),
    ("kernel_tilt_d", "double c = ax * ux + ay * uy + az * uz;", "double c = ax * ux + ay * uy - az * uz;"),
    ("orc_cone_dir", "double dh[3] = { sth * cos(phi), sth * sin(phi), cth };",
     "double dh[3] = { sth * sin(phi), sth * cos(phi), cth };"),
    ("orc_cone_dir", "double dh[3] = { sth * cos(phi), sth * sin(phi), cth };",
     "double dh[3] = { sth * cos(phi), -sth * sin(phi), cth };"),
    ("orc_cone_dir", "float thf = kernel36[6 * r + 0] * ORC_PI_F / 180.0f;",
     "float thf = kernel36[6 * r + 1] * ORC_PI_F / 180.0f;"),
    ("orc_pencil", "double da = dist - w + 0.7 * r80;", "double da = dist + w + 0.7 * r80;"),
    ("orc_pencil", "double da = dist - w + 0.7 * r80;", "double da = dist - w + 0.07 * r80;"),
    ("orc_pencil", "double da = dist - w + 0.7 * r80;", "double da = dist - w - 0.7 * r80;"),
    ("orc_pencil", "((double)c[0] * da + c[1]) * da + c[2];", "((double)c[1] * da + c[0]) * da + c[2];"),
    ("orc_pencil", "double sig_t = sig_air + sms;", "double sig_t = hypot(sig_air, sms);"),
    ("orc_pencil", "double sig_ht = hypot(sig_t, sig_h);", "double sig_ht = sig_t + sig_h;"),
    ("orc_pencil", "double sig_ht = hypot(sig_t, sig_h);", "double sig_ht = hypot(sig_t, hw);"),
    ("orc_pencil", "double H = om * idd / (2.0 * ORC_PI * sig_ht * sig_ht);",
     "double H = om * idd / (2.0 * ORC_PI * sig_t * sig_t);"),
    ("orc_pencil", "double P = (1.0 - om) * idd", "double P = (1.0 + om) * idd"),
    ("orc_pencil", "double ah = sig_ht != 0.0 ? -0.5 / (sig_ht * sig_ht)",
     "double ah = sig_ht != 0.0 ? -0.5 / (sig_t * sig_t)"),
    ("orc_pencil", "double dist = sqrt(dx * dx + dy * dy + dz * dz);",
     "double dist = sqrt(dx * dx + dy * dy);"),
    ("orc_pencil", "double d2 = rr - dp * dp / tt;", "double d2 = rr - dp * dp;"),
]


def _span(src, name):
    m = re.search(r"\n(?:static )?[a-z0-9_]+ \*?" + name + r"\(", src)
    assert m, name
    i = src.index("{", m.end())
    depth = 0
    for j in range(i, len(src)):
        depth += {"{": 1, "}": -1}.get(src[j], 0)
        if depth == 0:
            return m.start(), j + 1
    raise AssertionError(name)


def _mutate(src, func, old, new):
    a, b = _span(src, func)
    body = src[a:b]
    assert body.count(old) == 1, (func, old)
    return src[:a] + body.replace(old, new) + src[b:]


def _compile(csrc, path):
    c = path + ".c"
    with open(c, "w") as f:
        f.write(csrc)
    r = subprocess.run(["gcc", *CFLAGS, "-o", path, c, "-lm"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return path


def _checks(photon_model):
    out = [(f.__name__, f) for f in P.PINS]
    out.append(("pin_cone_directions", lambda: P.pin_cone_directions(photon_model.kernel)))
    out.append(("test_source_couch_rotates_about_y", T.test_source_couch_rotates_about_y))
    for ga, e in [(0.0, (0, -1, 0)), (90.0, (1, 0, 0))]:
        out.append((f"test_source_position_cardinal[{ga}]",
                    lambda ga=ga, e=e: T.test_source_position_cardinal(ga, e)))
    out.append(("test_pencil_lateral_integral_is_idd", T.test_pencil_lateral_integral_is_idd))
    out.append(("test_pencil_pure_gaussian_and_cax_distance",
                T.test_pencil_pure_gaussian_and_cax_distance))
    return out


def _failing(path, checks):
    failed = []
    with O.using(path):
        for name, fn in checks:
            try:
                fn()
            except AssertionError:
                failed.append(name)
    return failed


@pytest.fixture(scope="module")
def built(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("oracle_mutants"))
    src = open(SRC).read()
    jobs = [(src, os.path.join(d, "clean.so"))]
    for n, (func, old, new) in enumerate(MUTATIONS):
        jobs.append((_mutate(src, func, old, new), os.path.join(d, f"m{n:02d}.so")))
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        paths = list(ex.map(lambda j: _compile(*j), jobs))
    return paths[0], paths[1:]


def test_clean_copy_passes_every_check(built, photon_model):
    clean, _ = built
    assert _failing(clean, _checks(photon_model)) == []


@pytest.mark.parametrize("n", range(len(MUTATIONS)),
                         ids=[f"{m[0]}:{m[2][:40]}" for m in MUTATIONS])
def test_mutation_is_caught(built, photon_model, n):
    _, mutants = built
    failed = _failing(mutants[n], _checks(photon_model))
    assert failed, f"planted mistake not caught: {MUTATIONS[n]}"
