"""The constant tables of the FP64 exp and log of the step kernel against the
mathematics: every entry recomputed at 80 digits and rounded once to double."""
import decimal
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PHYS = os.path.join(ROOT, "swehr_amd", "csrc", "swehr_physics.h")
D = decimal.Decimal
decimal.getcontext().prec = 80


def _table(name):
    src = open(PHYS).read()
    body = re.search(name + r"\[[^=]*=\s*\{(.*?)\};", src, re.S).group(1)
    return [float.fromhex(x) for x in re.findall(r"-?0x[0-9a-fA-F.]+p[-+]?\d+", body)]


def _const(name):
    src = open(PHYS).read()
    return float.fromhex(re.search(name + r"\s*=\s*(-?0x[0-9a-fA-F.]+p[-+]?\d+)", src).group(1))


def test_exp2_fraction_table_is_correctly_rounded():
    tab = _table("kExp2Frac")
    assert len(tab) == 32
    ln2 = D(2).ln()
    for j, v in enumerate(tab):
        assert v == float((ln2 * j / 32).exp()), j


def test_log_table_reciprocals_and_split_logs():
    tab = _table("kLogTab")
    assert len(tab) == 3 * 23
    for n, j in enumerate(range(-9, 14)):
        F = D(1) + D(j) / 32
        inv, hi, lo = tab[3 * n: 3 * n + 3]
        L = F.ln()
        assert inv == float(1 / F), j
        assert hi == float(L), j
        assert lo == float(L - D(hi)), j


def test_ln2_split_head_is_exact_in_products():
    hi, lo = _const("kLn2Hi"), _const("kLn2Lo")
    mant = int.from_bytes(__import__("struct").pack(">d", hi), "big") & ((1 << 52) - 1)
    assert mant & ((1 << 11) - 1) == 0  # 42 significant bits: m * hi exact for |m| < 2^11
    assert abs(D(hi) + D(lo) - D(2).ln()) < D(10) ** -30
