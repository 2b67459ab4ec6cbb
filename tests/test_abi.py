"""The C-ABI library loads and exports every entry point its headers declare.

No compute call is made here (no GPU is needed)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADERS = sorted((ROOT / "include").rglob("*.h"))


def declared_functions():
    names = set()
    for h in HEADERS:
        text = "\n".join(l for l in h.read_text().splitlines() if not l.lstrip().startswith("#"))
        names.update(re.findall(r"(?:VL_EXPORT|VLHIP_API)[^;(]*?\b(\w+)\s*\(", text, re.S))
    return sorted(names)


def test_headers_declare_the_reference_api():
    names = set(declared_functions())
    for required in ["vl_gmm_new", "vl_gmm_cluster", "vl_gmm_get_posteriors", "vl_gmm_em",
                     "vl_get_gmm_data_posteriors_f", "vl_get_gmm_data_posteriors_d", "vl_fisher_encode",
                     "vl_vlad_encode", "vl_kmeans_quantize", "vl_kmeans_quantize_ANN", "vl_kmeans_quantize_ann",
                     "vl_kmeans_cluster", "vl_rand_seed", "vl_get_rand", "vl_set_num_threads",
                     "vlhip_fisher_encode_batch", "vlhip_gmm_posteriors"]:
        assert required in names, required


def test_library_exports_every_declared_symbol():
    import vlfeat_amd
    if not vlfeat_amd.LIB_PATH.exists():
        pytest.fail("libvl.so not built; run python -m vlfeat_amd.build")
    lib = ctypes.CDLL(str(vlfeat_amd.LIB_PATH), mode=ctypes.RTLD_LOCAL)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_kmeans_struct_layout_matches_reference_header():
    # VlKMeans is public in the reference (kmeans.h:45-68); inline accessors
    # compiled into clients read its fields by offset.
    from vlfeat_amd.capi import VlKMeans
    offs = {f: getattr(VlKMeans, f).offset for f, _ in VlKMeans._fields_}
    assert offs["dimension"] == 8 and offs["centers"] == 88 and offs["energy"] == 104
    assert ctypes.sizeof(VlKMeans) == 128  # gcc offsetof on vl/kmeans.h: 8, 88, 104, 128
