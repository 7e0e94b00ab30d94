"""Writes mwc1616_golden.json from the reference's own src/random.cpp (MWC1616 uniform and
Marsaglia polar normal), built by `make -C oracle ref` into oracle/_ref/ref_random.
Run in the container that holds the reference tree; the JSON is the committed fixture."""
import json
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle", "_ref", "ref_random")
COUNT = 200
out = {"generator": "reference src/random.cpp via oracle/_ref/ref_random", "count": COUNT, "seeds": []}
for a, b in ((42, 142), (1, 2), (123456789, 987654321)):
    lines = subprocess.run([EXE, str(a), str(b), str(COUNT)], check=True, capture_output=True, text=True).stdout.split()
    vals = [float(x) for x in lines[1::2]]
    out["seeds"].append({"a": a, "b": b, "uniform": [repr(v) for v in vals[:COUNT]],
                         "normal": [repr(v) for v in vals[COUNT:]]})
with open(os.path.join(HERE, "mwc1616_golden.json"), "w") as f:
    json.dump(out, f, indent=0)
print("wrote", COUNT, "draws x", len(out["seeds"]), "seeds")
