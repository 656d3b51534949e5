"""The south-building-24 real-data problem (24 views, 3,000 landmarks, 12,706 observations) from
`tests/golden/south_building_24.npz` as a reconstruction, and the reference's measured results on it
(`docs/experiments/gpu-budget/south-building-24-ba-{fixed,shared}.json`: 4 LM iterations,
`huberPx: 1`, the whole problem resident, the WebGPU tier)."""
from pathlib import Path

import numpy as np

FIXTURE = Path(__file__).resolve().parent / "golden" / "south_building_24.npz"

RECORD = {
    "fixed": {"cgIterations": 304, "acceptedSteps": 4, "initialRmsePx": 1.6565408120465435,
              "finalRmsePx": 0.1581431866127727},
    "shared": {"cgIterations": 317, "acceptedSteps": 4, "initialRmsePx": 1.6565408120465435,
               "finalRmsePx": 0.15806548731303538},
}
OPTIONS = {"iterations": 4, "huberPx": 1}


def reconstruction() -> dict:
    with np.load(FIXTURE) as f:
        z = {k: f[k] for k in f.files}
    views = [{"id": int(z["view_id"][i]),
              "camera": dict(zip(("fx", "fy", "cx", "cy", "k1", "k2"),
                                 (float(x) for x in z["camera"][i]))),
              "pose": {"R": z["R"][i].astype(np.float64), "t": z["t"][i].astype(np.float64)}}
             for i in range(len(z["view_id"]))]
    starts = z["obs_start"]
    landmarks = []
    for p in range(len(z["xyz"])):
        obs = [{"viewId": int(z["obs_view"][k]), "featureId": int(z["obs_feature"][k]),
                "pixel": [float(z["obs_pixel"][k, 0]), float(z["obs_pixel"][k, 1])]}
               for k in range(starts[p], starts[p + 1])]
        landmarks.append({"trackId": p, "xyz": z["xyz"][p].astype(np.float64),
                          "observations": obs})
    return {"views": views, "landmarks": landmarks}
