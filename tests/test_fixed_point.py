"""The wide fixed-point arithmetic of the Float64 sums (device_util.h:
fix_words, fixsq_words, sword_to_f64, var_words) against exact rational
arithmetic, on the host: every image is the exact truncation, every
conversion the correctly rounded value, and the variance finish the
correctly rounded (n*Q - T^2) / (n (n - ddof)) of its totals; and the
variance error bound the query compiler relies on holds for the truncated
totals. Builds tests/native/fixed_point_host.cpp with hipcc."""
import math
import os
import random
import shutil
import struct
import subprocess
from fractions import Fraction as F

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    if not (os.path.exists(HIPCC) or shutil.which("hipcc")):
        pytest.skip("hipcc is not installed")
    exe = str(tmp_path_factory.mktemp("fixed") / "fixed_point_host")
    subprocess.run([HIPCC, "-O1", "-std=c++17", "-o", exe, os.path.join(HERE, "native", "fixed_point_host.cpp")],
                   check=True, capture_output=True)
    p = subprocess.Popen([exe], stdin=subprocess.PIPE, stdout=subprocess.PIPE, text=True)

    def ask(line):
        p.stdin.write(line + "\n")
        p.stdin.flush()
        return p.stdout.readline().split()

    yield ask
    p.stdin.close()
    p.wait()


def bits(x):
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def dbl(b):
    return struct.unpack("<d", struct.pack("<Q", b))[0]


def words(v, n):
    v %= 1 << (64 * n)
    return [(v >> (64 * i)) & (2**64 - 1) for i in range(n)]


def trunc(fr):
    return int(fr) if fr >= 0 else -int(-fr)


def hexw(ws):
    return " ".join(f"{w:x}" for w in ws)


def test_images_and_conversions(harness):
    rnd = random.Random(1)
    for _ in range(2000):
        n = rnd.choice([2, 3, 4])
        x = rnd.choice([-1, 1]) * rnd.random() * 2.0 ** rnd.randint(-300, 300)
        if rnd.random() < 0.05:
            x = rnd.choice([5e-324, -1e-310, 2.5e-320])
        b = math.frexp(abs(x))[1] if x else 0
        s = 64 * n - 2 - b - rnd.randint(1, 40)
        assert [int(w, 16) for w in harness(f"fix {bits(x):x} {s} {n}")] == words(trunc(F(x) * F(2) ** s), n)
        s2 = 64 * n - 2 - 2 * b - rnd.randint(1, 40)
        assert [int(w, 16) for w in harness(f"sq {bits(x):x} {s2} {n}")] == words(int(F(x) * F(x) * F(2) ** s2), n)
        v = rnd.randint(-(2 ** (64 * n - 2)), 2 ** (64 * n - 2)) >> rnd.randint(0, 64 * n - 3)
        sc = rnd.randint(-200, 400)
        exact = F(v) / F(2) ** sc
        if abs(exact) < 2**1023:
            assert dbl(int(harness(f"tof {n} {sc} {hexw(words(v, n))}")[0], 16)) == float(exact)


def test_variance_finish_and_its_error_bound(harness):
    rnd = random.Random(2)
    for _ in range(300):
        nf, nq = rnd.choice([2, 3, 4]), rnd.choice([2, 3, 4])
        n = rnd.randint(2, 50)
        scale = 2.0 ** rnd.randint(-60, 60)
        vals = [rnd.gauss(0, 1) * scale * (10 ** rnd.randint(0, 8) if rnd.random() < 0.2 else 1)
                + rnd.choice([0, 1e6 * scale]) for _ in range(n)]
        b = max(math.frexp(abs(v))[1] for v in vals)
        L = max(1, n.bit_length())
        s, s2 = 64 * nf - 2 - b - L, 64 * nq - 2 - 2 * b - L
        T = sum(trunc(F(v) * F(2) ** s) for v in vals)
        Q = sum(int(F(v) * F(v) * F(2) ** s2) for v in vals)
        ddof = rnd.choice([0, 1])
        got = dbl(int(harness(f"var {nf} {s} {nq} {s2} {n} {ddof} {hexw(words(T, nf))} {hexw(words(Q, nq))}")[0], 16))
        num = max(F(0), F(n) * F(Q) / F(2) ** s2 - (F(T) / F(2) ** s) ** 2)
        assert got == float(num) / float(n * (n - ddof))
        # the compiler's bound: (ddof + 1)(2^-S2 + 2^(B+1-S) + 2^-2S) <= 2^(bitlen(ddof) + 2 + max(-S2, B+1-S))
        mean = sum(F(v) for v in vals) / n
        true = sum((F(v) - mean) ** 2 for v in vals) / (n - ddof)
        bound = F(2) ** (ddof.bit_length() + 2) * max(F(2) ** -s2, F(2) ** (b + 1 - s))
        assert abs(F(got) - true) <= bound + abs(true) * F(1, 2**50)


def test_two_word_variance_finish_and_256_bit_helpers(harness):
    # the finish of two-word totals (the usual width), and the 256-bit
    # shift and rounding it is made of, against exact rationals
    rnd = random.Random(3)
    for _ in range(400):
        n = rnd.randint(2, 10**6) if rnd.random() < 0.5 else rnd.randint(2, 50)
        b, L = rnd.randint(-40, 40), max(1, n.bit_length())
        s, s2 = 126 - b - L, 126 - 2 * b - L
        vals = [rnd.gauss(0, 1) * 2.0 ** (b - 1) for _ in range(min(n, 50))]
        T = sum(trunc(F(v) * F(2) ** s) for v in vals) * (n // len(vals))
        Q = sum(int(F(v) * F(v) * F(2) ** s2) for v in vals) * (n // len(vals))
        ddof = rnd.choice([0, 1])
        got = dbl(int(harness(f"var2 {s} {s2} {n} {ddof} {hexw(words(T, 2))} {hexw(words(Q, 2))}")[0], 16))
        num = max(F(0), F(n) * F(Q) / F(2) ** s2 - (F(T) / F(2) ** s) ** 2)
        assert got == float(num) / float(n * (n - ddof))
    for _ in range(2000):
        v = rnd.getrandbits(256) >> rnd.randint(0, 255)
        if rnd.random() < 0.1:
            v = rnd.choice([0, 1, 2**64, 2**128 - 1, 2**192, 2**255 + 7])
        sc = rnd.randint(-100, 900)
        exact = F(v) / F(2) ** sc
        if exact < 2**1023:
            assert dbl(int(harness(f"u256 {sc} {hexw(words(v, 4))}")[0], 16)) == float(exact)
        sh = rnd.randint(0, 255)
        assert [int(w, 16) for w in harness(f"shl {sh} {hexw(words(v, 4))}")] == words(v << sh, 4)
