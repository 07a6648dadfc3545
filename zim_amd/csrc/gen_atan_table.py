"""Writes atan_table.inc: atan(i/64), i = 0..64, rounded to double.

The values come from numpy's extended-precision (x87 80-bit) arctan, whose
64-bit significand rounds to the nearest double except in double-rounding
ties, which the table test in tests/ checks for. Run from this directory:
    python gen_atan_table.py > atan_table.inc
"""
import numpy as np

assert np.finfo(np.longdouble).nmant >= 63, "needs x87 extended precision"
print("/* atan(i/64) for i = 0..64, generated by gen_atan_table.py */")
for i in range(65):
    v = np.arctan(np.longdouble(i) / np.longdouble(64))
    print(f"    {float(v).hex()},")
