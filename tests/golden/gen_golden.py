"""Writes tests/golden/*.json: known-answer vectors computed with pure-Python
big integers, independent of the C oracle and of the HIP library.

  omega.json   : ark-convention roots of unity w_N = g^((p-1)/N) for every
                 N = 2^k up to each field's two-adicity (src/fields.rs:3-17,
                 src/encoder.rs:34-35), checked primitive here.
  dft.json     : brute-force DFT out[j] = sum_i x[i] w^(ij) (the closed form
                 the reference's Dft encoder states, src/encoders/dft.rs:9-22)
                 for N = 2..256 per field on seeded inputs, plus the inverse.
  lde.json     : coset LDE by direct polynomial evaluation at shift * w_M^i
                 and its decode, for small n, blowup 4.
  fieldops.json: add/sub/mul edge cases (0, 1, p-1, 2^32 +- 1, high words).

Run:  python tests/golden/gen_golden.py
"""
import json
import os
import random

HERE = os.path.dirname(os.path.abspath(__file__))
FIELDS = [("babybear", 2013265921, 31), ("koalabear", 2130706433, 3),
          ("goldilocks", 2**64 - 2**32 + 1, 7), ("teddybear", 3221225473, 5)]


def two_adicity(p):
    s, t = 0, p - 1
    while t % 2 == 0:
        t //= 2
        s += 1
    return s, t


def omega(p, g, n):
    return pow(g, (p - 1) // n, p)


def dft(x, w, p):
    n = len(x)
    return [sum(x[i] * pow(w, i * j, p) for i in range(n)) % p for j in range(n)]


def main():
    rng = random.Random(20241203)
    om, dv, lde, fo = {}, {}, {}, {}
    for name, p, g in FIELDS:
        s, t = two_adicity(p)
        roots = {}
        for k in range(0, s + 1):
            w = omega(p, g, 1 << k)
            assert pow(w, 1 << k, p) == 1
            if k:
                assert pow(w, 1 << (k - 1), p) != 1
            roots[str(k)] = w
        om[name] = {"p": p, "g": g, "two_adicity": s, "omega": roots}

        cases = []
        for k in range(1, 9):
            n = 1 << k
            x = [rng.randrange(p) for _ in range(n)]
            w = omega(p, g, n)
            X = dft(x, w, p)
            winv = pow(w, p - 2, p)
            ninv = pow(n, p - 2, p)
            back = [(v * ninv) % p for v in dft(X, winv, p)]
            assert back == x
            cases.append({"n": n, "x": x, "X": X})
        # edge inputs: delta, constant, all p-1
        for n in (16, 64):
            w = omega(p, g, n)
            for label, x in (("delta", [1] + [0] * (n - 1)), ("const", [5] * n), ("pm1", [p - 1] * n)):
                cases.append({"n": n, "x": x, "X": dft(x, w, p), "label": label})
        dv[name] = cases

        lcases = []
        for n in (4, 16, 64):
            blow = 4
            m = n * blow
            c = [rng.randrange(p) for _ in range(n)]
            wm = omega(p, g, m)
            ev = []
            for i in range(m):
                pt = g * pow(wm, i, p) % p
                acc, pw = 0, 1
                for cj in c:
                    acc = (acc + cj * pw) % p
                    pw = pw * pt % p
                ev.append(acc)
            lcases.append({"n": n, "blowup": blow, "shift": g, "coeffs": c, "evals": ev})
        lde[name] = lcases

        edges = [0, 1, 2, p - 1, p - 2, (p - 1) // 2, (p + 1) // 2]
        if p > 2**32:
            edges += [2**32 - 1, 2**32, 2**32 + 1, 2**63, 2**64 - 2**33 + 5, p - 2**32]
        else:
            edges += [2**31 - 1 if 2**31 - 1 < p else 5, p - 2**16]
        a = [e for e in edges for _ in edges] + [rng.randrange(p) for _ in range(200)]
        b = [e for _ in edges for e in edges] + [rng.randrange(p) for _ in range(200)]
        fo[name] = {"a": a, "b": b,
                    "add": [(x + y) % p for x, y in zip(a, b)],
                    "sub": [(x - y) % p for x, y in zip(a, b)],
                    "mul": [(x * y) % p for x, y in zip(a, b)]}

    for fname, obj in (("omega.json", om), ("dft.json", dv), ("lde.json", lde), ("fieldops.json", fo)):
        with open(os.path.join(HERE, fname), "w") as fh:
            json.dump(obj, fh, separators=(",", ":"))
        print("wrote", fname)


if __name__ == "__main__":
    main()
