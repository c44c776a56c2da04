"""Write splitmix64.json: the first 8 SplitMix64 outputs of every workload seed 0xD1FF0000 + c.

The values come from the published recurrence (Steele, Lea & Flood 2014):
  z = (state += 0x9E3779B97F4A7C15); z = (z ^ z>>30) * 0xBF58476D1CE4E5B9;
  z = (z ^ z>>27) * 0x94D049BB133111EB; return z ^ z>>31
computed here with Python integers, independent of both numpy and the C implementations.
"""
import json
import os

M = (1 << 64) - 1


def splitmix64(seed, n):
    s = seed
    out = []
    for _ in range(n):
        s = (s + 0x9E3779B97F4A7C15) & M
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        out.append(z ^ (z >> 31))
    return out


if __name__ == "__main__":
    seeds = {str(c): [str(v) for v in splitmix64(0xD1FF0000 + c, 8)] for c in range(5)}
    path = os.path.join(os.path.dirname(__file__), "splitmix64.json")
    json.dump({"generator": "SplitMix64", "u01": "(z >> 11) * 2^-53", "first_outputs": seeds}, open(path, "w"),
              indent=1)
    print("wrote", path)
