"""Writes the table of log(k!) for k < 1024, each entry the correctly rounded
double of the exact value (computed with 50-digit decimal arithmetic), as a C
initializer list. Used for the engine (stg_lfact_table.inc) and the CPU oracle
(oracle/lfact_table.inc):

    python steps_amd/csrc/gen_lfact_table.py steps_amd/csrc/stg_lfact_table.inc
    python steps_amd/csrc/gen_lfact_table.py oracle/lfact_table.inc
"""
import decimal
import sys

N = 1024


def table():
    decimal.getcontext().prec = 50
    acc = decimal.Decimal(0)
    out = [0.0]
    for k in range(1, N):
        acc += decimal.Decimal(k).ln()
        out.append(float(acc))
    return out


def main(path):
    vals = table()
    with open(path, "w") as f:
        f.write("/* log(k!) for k = 0..%d, correctly rounded; generated by "
                "steps_amd/csrc/gen_lfact_table.py */\n" % (N - 1))
        for i in range(0, N, 4):
            f.write("    " + ", ".join(v.hex() for v in vals[i:i + 4]) + ",\n")


if __name__ == "__main__":
    main(sys.argv[1])
