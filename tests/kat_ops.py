"""Expand the op descriptions of tests/golden/reference_kats.json."""
import math

OPS_NAMES = {"H": 0, "X": 1, "Y": 2, "Z": 3, "P": 4, "RX": 5, "RY": 6, "RZ": 7, "U": 8}



def kat_op_dicts(case):
    n = case["n"]
    spec = case["ops"]
    out = []

    def o(kind, target, theta=0.0, phi=0.0, lam=0.0, ones=0):
        return {"kind": kind, "target": target, "theta": theta, "phi": phi, "lam": lam,
                "t1": 0, "ones": ones, "zeros": 0, "classical": 0}

    if spec == "X_all":
        return [o(1, q) for q in range(n)]
    if spec == "Y_all":
        return [o(2, q) for q in range(n)]
    if spec == "qcbm":
        pairs = [(i, (i + 1) % n) for i in range(n)][: n - 1]
        for i in range(n):
            out += [o(5, i, 1.0), o(7, i, 1.0)]
        out += [o(1, b, ones=1 << a) for a, b in pairs]
        for _ in range(9):
            for i in range(n):
                out += [o(7, i, 1.0), o(5, i, 1.0), o(7, i, 1.0)]
            out += [o(1, b, ones=1 << a) for a, b in pairs]
        for i in range(n):
            out += [o(7, i, 1.0), o(5, i, 1.0)]
        return out
    if spec == "value_encoding_v4":
        v = 4.0
        for t in range(n):
            out.append(o(0, t))
        for t in range(n):
            out.append(o(4, t, 2.0 * math.pi / float(2 ** (t + 1)) * v))
        targets = list(range(n - 1, -1, -1))
        for j in range(len(targets) - 1, -1, -1):
            out.append(o(0, targets[j]))
            for k in range(j - 1, -1, -1):
                out.append(o(4, targets[k], -math.pi / float(2 ** (j - k)), ones=1 << targets[j]))
        return out
    for item in spec:
        name = item[0]
        if name == "CH":
            out.append(o(0, item[2], ones=1 << item[1]))
        elif name == "CRZ":
            out.append(o(7, item[2], item[3], ones=1 << item[1]))
        elif name == "CU":
            out.append(o(8, item[2], item[3], item[4], item[5], ones=1 << item[1]))
        else:
            k = OPS_NAMES[name]
            args = item[2:]
            out.append(o(k, item[1], *args))
    return out


def expected_vectors(case):
    """(indices, re, im) lists of the amplitudes a case pins."""
    N = 1 << case["n"]
    if "all_re" in case:
        return list(range(N)), [case["all_re"]] * N, [case["all_im"]] * N
    amps = case["amps"]
    if case.get("only_listed"):
        return [a[0] for a in amps], [a[1] for a in amps], [a[2] for a in amps]
    re = [0.0] * N
    im = [0.0] * N
    for i, r, m in amps:
        re[i], im[i] = r, m
    return list(range(N)), re, im
