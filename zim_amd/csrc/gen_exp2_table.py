"""Writes exp2_table.inc: 2^(j/256), j = 0..255, each the double nearest to
the exact value.

The values are computed with Python's decimal module at 60 significant digits
and rounded to double by float() of the decimal string (correctly rounded in
CPython); every fourth entry must equal the 64-entry table kExp2Tab64 of
psgc_math.h, which the table test in tests/ checks. Run from this directory:
    python gen_exp2_table.py > exp2_table.inc
"""
from decimal import Decimal, getcontext

getcontext().prec = 60
two = Decimal(2)
ln2 = two.ln()
print("/* 2^(j/256) for j = 0..255, generated by gen_exp2_table.py */")
for j in range(256):
    v = (ln2 * j / 256).exp()
    print(f"    {float(str(v)).hex()},")
