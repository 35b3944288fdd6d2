"""Embed the period-kernel headers in the engine library as C++ string
constants, for the model-specialised kernels it compiles at run time with
hipRTC. Build step: python3 gen_rtc_sources.py > stg_rtc_sources.inc"""
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
FILES = [("kRtcKernels", "stg_kernels.hpp"), ("kRtcRandom", "stg_random.hpp"), ("kRtcLfact", "stg_lfact_table.inc")]
CHUNK = 16000  # characters per raw string literal, split on line ends

out = ["// Generated by gen_rtc_sources.py from the period-kernel headers. Do not edit."]
for sym, name in FILES:
    text = (HERE / name).read_text()
    assert ")STGRTC\"" not in text
    parts, cur = [], ""
    for line in text.splitlines(keepends=True):
        if len(cur) + len(line) > CHUNK and cur:
            parts.append(cur)
            cur = ""
        cur += line
    parts.append(cur)
    out.append(f"static const char *const {sym}[] = {{")
    for p in parts:
        out.append(f'R"STGRTC({p})STGRTC",')
    out.append("};")
sys.stdout.write("\n".join(out) + "\n")
