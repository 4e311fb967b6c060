"""Writes tests/golden/reference_native.json: the inputs and expected outputs the reference's own golden files hold
for the path engine (tests/golden/{calibration_simulated,native_cpu,native_gpu,labjs}.json of the reference), cut
down to what the tests here read. Data only; run in the development container, where the reference is mounted:

    python tests/golden/make_fixtures.py [/root/reference]
"""

from __future__ import annotations

import json
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent


def main(ref_root: str = "/root/reference") -> None:
    gold = Path(ref_root) / "tests" / "golden"
    cal = json.loads((gold / "calibration_simulated.json").read_text())["values"]
    cpu = json.loads((gold / "native_cpu.json").read_text())["values"]
    gpu = json.loads((gold / "native_gpu.json").read_text())["values"]
    lab = json.loads((gold / "labjs.json").read_text())["values"]["values"]
    out = {
        "source": "reference tests/golden: calibration_simulated.json, native_cpu.json, native_gpu.json, labjs.json",
        "simulated_context": {
            "market": {"spot": 100.0, "r": 0.04, "q": 0.012},
            "surface": cal["surface"],
            "heston": cal["heston"],
        },
        "native_cpu": {
            "paths": {k: cpu["paths"][k] for k in ("bs", "lv", "heston", "slv", "basket_localcorr")},
            "leverage": cpu["leverage"],
            "fixed": cpu["fixed"],
        },
        "native_gpu": {
            "paths": {k: gpu["paths"][k] for k in ("bs", "lv", "heston", "slv")},
            "leverage": {k: gpu["leverage"][k] for k in ("x_grids", "values")},
            "fixed": gpu["fixed"],
        },
        "labjs": {
            "heston_params": {"v0": 0.0228466, "kappa": 3.0892636, "theta": 0.0537246, "xi": 1.3077452,
                              "rho": -0.6879607},
            "heston_fourier_cases": [[100, 1.0, "call"], [119, 1.0, "call"], [85, 0.5, "put"], [110, 2.0, "put"]],
            "heston_fourier": lab["heston_fourier"],
            "heston_qe_call_119": lab["heston_qe_barrier"],
            "geo_asian_cases": [[100, 100, 1.0, 0.04, 0.01, 0.20, "call"], [100, 90, 0.5, 0.03, 0.0, 0.35, "put"],
                                [100, 120, 2.0, 0.05, 0.02, 0.15, "call"], [100, 105, 0.25, 0.0, 0.01, 0.25, "put"]],
            "geo_asian": lab["geo_asian"],
        },
    }
    (HERE / "reference_native.json").write_text(json.dumps(out, indent=1))
    print(HERE / "reference_native.json")


if __name__ == "__main__":
    main(*sys.argv[1:])
