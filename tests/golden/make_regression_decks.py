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


def vsg_deck():
    """tests/regression/dic_challenge_12_sorted (a region of interest with an excluded circle, step 35,
    USE_NEIGHBOR_VALUES_FIRST_STEP_ONLY, normal strains, the VSG strain post-processor with a 360-px
    window, sorted comma-separated output without row ids), cut to its first two frames: the reference
    image oht_cfrp_00.tiff and the deformed oht_cfrp_01/02.tiff (8-bit grey palette TIFFs whose palette is
    the identity, stored as their pixel values in images.npz) and the golds DICe_solution_01/02.txt."""
    import numpy as np
    from PIL import Image
    name = "dic_challenge_12_sorted"
    src, dst = os.path.join(SRC, name), os.path.join(DST, name)
    os.makedirs(os.path.join(dst, "gold"), exist_ok=True)
    inp = top_params(ET.parse(os.path.join(src, "input.xml")).getroot())
    cor_root = ET.parse(os.path.join(src, inp["correlation_parameters_file"])).getroot()
    cor = top_params(cor_root)
    frames = {}
    for k in range(3):
        im = Image.open(os.path.join(src, "images", "oht_cfrp_%02d.tiff" % k))
        pal = im.getpalette()
        assert all(pal[3 * i] == pal[3 * i + 1] == pal[3 * i + 2] == i for i in range(256))
        frames["frame_%d" % k] = np.asarray(im, dtype=np.uint8)
    np.savez_compressed(os.path.join(dst, "images.npz"), **frames)
    shutil.copy(os.path.join(src, inp["subset_file"]), os.path.join(dst, "subsets.txt"))
    spec = cor_root.find("ParameterList[@name='output_spec']")
    fields = sorted(spec.findall("Parameter"), key=lambda q: int(q.get("value")))
    deck = {"subset_size": int(inp["subset_size"]), "step_size": int(inp["step_size"]), "separate": False,
            "prefix": inp.get("output_prefix", "DICe_solution"),
            "reference_image": inp["image_folder"] + "oht_cfrp_00.tiff",
            "deformed_image": inp["image_folder"] + "oht_cfrp_01.tiff",
            "deformed_images": [inp["image_folder"] + "oht_cfrp_%02d.tiff" % k for k in (1, 2)],
            "frames": 2, "first_frame_index": int(inp["start_image_index"]),
            "num_frames": int(inp["end_image_index"]) - int(inp["start_image_index"]) + 1,
            "subset_file": "subsets.txt", "images": "images.npz",
            "vsg_window": int(cor_root.find("ParameterList[@name='post_process_vsg_strain']")
                              .find("Parameter").get("value")),
            "sort": cor.get("sort_txt_output") == "true", "omit_row_id": cor.get("omit_output_row_id") == "true",
            "delimiter": cor.get("output_delimiter", " "),
            "fields": [q.get("name") for q in fields],
            "gold": ["DICe_solution_01.txt", "DICe_solution_02.txt"],
            "diff_tolerance": 0.05, "diff_floor": 5.0e-5}
    for k in ("enable_rotation", "enable_normal_strain", "enable_shear_strain"):
        deck[k] = 1 if cor.get(k) == "true" else 0
    for key, names in ENUMS.items():
        if key in cor:
            deck[key] = names.index(cor[key].upper())
    for g in deck["gold"]:
        shutil.copy(os.path.join(src, "gold", g), os.path.join(dst, "gold", g))
    for root, _, files in os.walk(dst):
        for f in files:
            os.chmod(os.path.join(root, f), 0o644)
    with open(os.path.join(dst, "deck.json"), "w") as f:
        json.dump(deck, f, indent=1)


def unicode_deck():
    """tests/regression/unicode_filename: the image ../../component/images/\u00e6b\u00e4rn\u0243 0.tif (8-bit grey,
    uncompressed) correlated against itself on the regular grid of 15-px subsets at step 50 (GRADIENT_BASED,
    all six parameters, USE_FIELD_VALUES), comma-separated output without row ids; the image is stored as its
    pixel values in images.npz and the test writes it back under its non-ASCII name."""
    import unicodedata

    import numpy as np
    from PIL import Image
    name = "unicode_filename"
    src, dst = os.path.join(SRC, name), os.path.join(DST, name)
    os.makedirs(os.path.join(dst, "gold"), exist_ok=True)
    inp = top_params(ET.parse(os.path.join(src, "input.xml")).getroot())
    cor_root = ET.parse(os.path.join(src, inp["correlation_parameters_file"])).getroot()
    cor = top_params(cor_root)
    image = inp["image_file_prefix"] + inp["reference_image_index"] + inp["image_file_extension"]
    # input.xml spells the prefix in decomposed form (a + U+0308), as the gold's header prints it; the file on
    # disk has the composed form
    im = Image.open(os.path.join(src, inp["image_folder"], unicodedata.normalize("NFC", image)))
    assert im.mode == "L"
    np.savez_compressed(os.path.join(dst, "images.npz"), frame_0=np.asarray(im, dtype=np.uint8))
    spec = cor_root.find("ParameterList[@name='output_spec']")
    fields = sorted(spec.findall("Parameter"), key=lambda q: int(q.get("value")))
    deck = {"subset_size": int(inp["subset_size"]), "step_size": int(inp["step_size"]), "separate": False,
            "prefix": inp.get("output_prefix", "DICe_solution"), "image_file": image,
            "reference_image": inp["image_folder"] + image, "deformed_image": inp["image_folder"] + image,
            "deformed_images": [inp["image_folder"] + image], "frames": 1, "images": "images.npz",
            "omit_row_id": cor.get("omit_output_row_id") == "true", "delimiter": cor.get("output_delimiter", " "),
            "fields": [q.get("name") for q in fields], "gold": ["DICe_solution_0.txt"],
            "diff_tolerance": 0.05, "diff_floor": 5.0e-5}
    for k in ("enable_rotation", "enable_normal_strain", "enable_shear_strain"):
        deck[k] = 1 if cor.get(k) == "true" else 0
    for key, names in ENUMS.items():
        if key in cor:
            deck[key] = names.index(cor[key].upper())
    for g in deck["gold"]:
        shutil.copy(os.path.join(src, "gold", g), os.path.join(dst, "gold", g))
    for root, _, files in os.walk(dst):
        for f in files:
            os.chmod(os.path.join(root, f), 0o644)
    with open(os.path.join(dst, "deck.json"), "w", encoding="utf-8") as f:
        json.dump(deck, f, indent=1, ensure_ascii=False)


if __name__ == "__main__":
    main()
    vsg_deck()
    unicode_deck()
