"""Pattern strings used by the tests (reference strings verbatim)."""
import os

TRIANGLE = "011101110"
K4 = "0111101111011110"  # scripts/run_pattern_matching.sh:15 (P1)
K5 = "0111110111110111110111110"  # scripts/run_pattern_matching.sh:24
HOUSE = "0100110110010110110010100"  # scripts/run_pattern_matching.sh:17 (P2)
C5 = "0100110100010100010110010"
P3 = "011011101110110101011000110000101000"  # scripts/run_pattern_matching.sh:19 (P3)
P4 = "0111111101111111011101110100111100011100001100000"  # scripts/run_pattern_matching.sh:21 (P4)

HERE = os.path.dirname(os.path.abspath(__file__))


def connected_patterns(n):
    """All connected graphs on n vertices up to isomorphism (fixture file)."""
    path = os.path.join(HERE, "golden", "connected_patterns.txt")
    out = []
    with open(path) as f:
        for line in f:
            k, s = line.split()
            if int(k) == n:
                out.append(s)
    return out


def clique(k):
    return "".join("0" if i == j else "1" for i in range(k) for j in range(k))
