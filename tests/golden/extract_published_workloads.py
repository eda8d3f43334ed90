"""Writes tests/golden/published_workloads.json from the skbench result bundles.

Run in the development container only (it reads the reference checkout at
/root/reference, which does not exist on GPU machines):

    python tests/golden/extract_published_workloads.py

The output is pure data: for every distinct published workload, the
retained-mode count, Nj, the FNV-1a permutation hashes
(workload.permutationHashes) and, for the synthetic K^2 family, the
vertical group count, group size statistics and group-order hash
(workload.grouping).  Entries that disagree with each other for the same
workload are dropped (some bundles come from experiments with another
retention rule).
"""
from __future__ import annotations

import glob
import json
import os
import sys

REFERENCE = "/root/reference/results/published"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "published_workloads.json")
K2_FAMILY = "synthetic-k2-grouped-orthonormal-dct2-pair-rotation-v1"


def documents(path):
    text = open(path, encoding="utf-8").read()
    try:
        yield json.loads(text)
        return
    except ValueError:
        pass
    for line in text.splitlines():
        line = line.strip().rstrip(",")
        if line.startswith("{"):
            try:
                yield json.loads(line)
            except ValueError:
                continue


def main() -> int:
    if not os.path.isdir(REFERENCE):
        print("reference results not present; nothing to do", file=sys.stderr)
        return 1
    table: dict[str, dict] = {}
    conflicts: set[str] = set()
    for path in sorted(glob.glob(os.path.join(REFERENCE, "**", "*.json"), recursive=True)):
        for doc in documents(path):
            w = doc.get("workload") if isinstance(doc, dict) else None
            if not isinstance(w, dict) or "Nx" not in w:
                continue
            aa = w.get("antialias", {})
            if aa.get("horizontalPolicy") != "radial-two-thirds" or not aa.get("enabled", False):
                continue
            hashes = w.get("permutationHashes", {})
            mode_hash = hashes.get("retainedModeOrder", "")
            if not mode_hash.startswith("fnv1a64:"):
                continue
            # bundles without a WVM full-spectrum order come from side
            # experiments with another mode enumeration (FFTW++ convolution)
            if not hashes.get("wvmFullSpectrumOrder", "").startswith("fnv1a64:"):
                continue
            grouping = w.get("grouping", {})
            family = grouping.get("verticalMatrixFamilyId")
            if family not in ("orthonormal-dct2-truncated-v1", K2_FAMILY):
                continue
            key = "%dx%dx%dx%d" % (w["Nx"], w["Ny"], w["Nz"], w["fields"])
            entry = {
                "nx": w["Nx"],
                "ny": w["Ny"],
                "nz": w["Nz"],
                "fields": w["fields"],
                "planes": w.get("planes"),
                "nkl": w.get("Nkl"),
                "nj": w.get("Nj"),
                "halfRows": w.get("H"),
                "retainedModeOrder": mode_hash,
            }
            entry["wvmFullSpectrumOrder"] = hashes["wvmFullSpectrumOrder"]
            nbytes = w.get("bytes", {})
            for name in ("real", "fullSpectrum", "retainedSpectrum", "modalSpectrum"):
                if nbytes.get(name):
                    entry["bytes_" + name] = nbytes[name]
            if family == K2_FAMILY and grouping.get("verticalGroupOrderHash"):
                entry["k2GroupCount"] = grouping.get("verticalGroupCount")
                entry["k2GroupModes"] = grouping.get("verticalGroupModes")
                entry["k2GroupColumns"] = grouping.get("verticalGroupColumns")
                entry["k2GroupOrderHash"] = grouping.get("verticalGroupOrderHash")
            entry["source"] = os.path.relpath(path, os.path.dirname(REFERENCE))
            old = table.get(key)
            if old is None:
                table[key] = entry
                continue
            for field, value in entry.items():
                if field == "source":
                    continue
                if field in old and old[field] != value:
                    conflicts.add(key + ":" + field)
                elif field not in old:
                    old[field] = value
    for item in sorted(conflicts):
        key, field = item.split(":", 1)
        if key in table:
            table[key].pop(field, None)
            table[key].setdefault("dropped", []).append(field)
    with open(OUT, "w", encoding="utf-8") as handle:
        json.dump(dict(sorted(table.items())), handle, indent=1, sort_keys=True)
        handle.write("\n")
    print("wrote %d workloads, %d conflicting fields dropped" % (len(table), len(conflicts)))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
