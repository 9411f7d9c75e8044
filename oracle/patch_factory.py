"""Applies the "hip" branch of INTEGRATION.md to the reference's backend
factory (src/libmpcd/common/backend_factory.hpp:15-27) and writes the result
to a build directory (test infrastructure: used by oracle/Makefile to build
the reference's own pybind11 module with the hip backend registered).

    python patch_factory.py <reference backend_factory.hpp> <output path>

Two insertions, nothing else changes:
  * after the USE_CUDA include block: #include "hip_backend/hip_backend.hpp"
    under USE_HIP;
  * before the final else that throws "Unsupported backend type": a branch
    that returns std::make_unique<hip::HipBackend>(params) for "hip".
Fails if either anchor is missing.
"""
from __future__ import annotations

import re
import sys
from pathlib import Path

INCLUDE = '#ifdef USE_HIP\n    #include "hip_backend/hip_backend.hpp"\n#endif\n'
BRANCH = ('        #ifdef USE_HIP\n'
          '        else if (backend_type == "hip") {\n'
          '            return std::make_unique<hip::HipBackend>(params);\n'
          '        }\n'
          '        #endif\n')


def patch(text: str) -> str:
    inc = re.search(r'#ifdef USE_CUDA\s*\n\s*#include "cuda_backend/cuda_backend.hpp"\s*\n#endif\s*\n', text)
    if not inc:
        raise SystemExit("anchor not found: the USE_CUDA include block")
    text = text[:inc.end()] + INCLUDE + text[inc.end():]
    thr = re.search(r'\n([ \t]*else\s*\{\s*\n\s*throw std::runtime_error\("Unsupported backend type"\);)', text)
    if not thr:
        raise SystemExit('anchor not found: the else branch throwing "Unsupported backend type"')
    return text[:thr.start() + 1] + BRANCH + text[thr.start() + 1:]


def main(argv):
    src, dst = Path(argv[1]), Path(argv[2])
    dst.parent.mkdir(parents=True, exist_ok=True)
    dst.write_text(patch(src.read_text()))


if __name__ == "__main__":
    main(sys.argv)
