"""Fixtures of three decks of the reference's regression suite (tests/regression/<deck>):
their subset files and gold outputs (data the reference's own tests hold), plus the deck
parameters this path needs, read from each deck's input.xml / params.xml into deck.json.

Run from the repository root: python tests/golden/make_regression_decks.py
(the images of these decks, ../../component/images/{ref,def}Speckled.tif, are the speckle
pair already in speckle_512.npz; defSpeckledCopy.tif, the second frame of
parallel_seed_first_step_only, is byte for byte defSpeckled.tif)."""
import json
import os
import shutil
import xml.etree.ElementTree as ET

SRC = "/root/reference/tests/regression"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "regression_decks")
DECKS = ["conformal_subset_def", "single_speckled_shift", "parallel_subsets", "parallel_subsets_seed",
         "parallel_seed_first_step_only"]
ENUMS = {"interpolation_method": ["BILINEAR", "BICUBIC", "KEYS_FOURTH"],
         "optimization_method": ["SIMPLEX", "GRADIENT_BASED", "GRADIENT_THEN_SEARCH", "SIMPLEX_THEN_GRADIENT_BASED",
                                 "GRADIENT_BASED_THEN_SIMPLEX"],
         "correlation_routine": ["GENERIC_ROUTINE", "TRACKING_ROUTINE"],
         "initialization_method": ["USE_FIELD_VALUES", "USE_NEIGHBOR_VALUES", "USE_NEIGHBOR_VALUES_FIRST_STEP_ONLY"]}


def top_params(root):
    return {p.get("name").lower(): p.get("value") for p in root.findall("Parameter")}


def main():
    for name in DECKS:
        src, dst = os.path.join(SRC, name), os.path.join(DST, name)
        os.makedirs(os.path.join(dst, "gold"), exist_ok=True)
        inp_root = ET.parse(os.path.join(src, "input.xml")).getroot()
        inp = top_params(inp_root)
        cor_root = ET.parse(os.path.join(src, inp["correlation_parameters_file"])).getroot()
        cor = top_params(cor_root)
        deck = {"subset_size": int(inp["subset_size"]),
                "separate": inp.get("separate_output_file_for_each_subset", "false") == "true",
                "prefix": inp.get("output_prefix", "DICe_solution"),
                "reference_image": inp["image_folder"] + inp["reference_image"]}
        defs = inp_root.find("ParameterList[@name='deformed_images']")
        deck["deformed_image"] = inp["image_folder"] + defs.find("Parameter").get("name")
        deck["frames"] = len(defs.findall("Parameter"))
        deck["deformed_images"] = [inp["image_folder"] + q.get("name") for q in defs.findall("Parameter")]
        if "subset_file" in inp:
            shutil.copy(os.path.join(src, inp["subset_file"]), os.path.join(dst, "subsets.txt"))
            deck["subset_file"] = "subsets.txt"
        if "step_size" in inp:
            deck["step_size"] = int(inp["step_size"])
        for key, names in ENUMS.items():
            if key in cor:
                deck[key] = names.index(cor[key].upper())
        spec = cor_root.find("ParameterList[@name='output_spec']")
        if spec is not None:
            deck["fields"] = [p.get("name") for p in spec.findall("Parameter") if p.get("value") == "true"]
        deck["gold"] = sorted(os.listdir(os.path.join(src, "gold")))
        for g in deck["gold"]:
            shutil.copy(os.path.join(src, "gold", g), os.path.join(dst, "gold", g))
        # the comparison of the reference's regression suite (tests/regression/CMakeLists.txt)
        deck["diff_tolerance"] = 0.05
        deck["diff_floor"] = 5.0e-5
        for root, _, files in os.walk(dst):
            for f in files:
                os.chmod(os.path.join(root, f), 0o644)
        with open(os.path.join(dst, "deck.json"), "w") as f:
            json.dump(deck, f, indent=1)


if __name__ == "__main__":
    main()
