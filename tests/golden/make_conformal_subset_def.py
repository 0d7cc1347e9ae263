"""Fixture of the reference's regression deck tests/regression/conformal_subset_def:
its subset file and gold outputs (data the reference's own test holds), plus the
deck parameters this path needs, read from the deck's input.xml / params.xml.

Run from the repository root: python tests/golden/make_conformal_subset_def.py
(the images of the deck are the speckle pair already in speckle_512.npz)."""
import json
import os
import shutil
import xml.etree.ElementTree as ET

SRC = "/root/reference/tests/regression/conformal_subset_def"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "conformal_subset_def")


def params(path):
    out = {}
    for p in ET.parse(path).getroot().iter("Parameter"):
        out[p.get("name")] = p.get("value")
    return out


def main():
    os.makedirs(os.path.join(DST, "gold"), exist_ok=True)
    inp = params(os.path.join(SRC, "input.xml"))
    cor = params(os.path.join(SRC, "params.xml"))
    shutil.copy(os.path.join(SRC, inp["subset_file"]), os.path.join(DST, "subset_defs.txt"))
    gold = sorted(os.listdir(os.path.join(SRC, "gold")))
    for g in gold:
        shutil.copy(os.path.join(SRC, "gold", g), os.path.join(DST, "gold", g))
    deformed = [p.get("name") for p in ET.parse(os.path.join(SRC, "input.xml")).getroot().iter("Parameter")
                if p.get("type") == "bool" and p.get("name").endswith(".tif")]
    deck = {
        "subset_size": int(inp["subset_size"]),
        "separate_output_file_for_each_subset": inp["separate_output_file_for_each_subset"] == "true",
        "reference_image": inp["image_folder"] + inp["reference_image"],
        "deformed_image": inp["image_folder"] + deformed[0],
        "correlation_routine": cor.get("correlation_routine", "GENERIC_ROUTINE"),
        "gold": gold,
        # the comparison of the reference's regression suite (tests/regression/CMakeLists.txt)
        "diff_tolerance": 0.05,
        "diff_floor": 5.0e-5,
    }
    with open(os.path.join(DST, "deck.json"), "w") as f:
        json.dump(deck, f, indent=1)


if __name__ == "__main__":
    main()
