"""Writes launch_layout.json: the `layout` object of the reference's launch
config fixture (framework/tests/launch_config.json), the data the bundle's
layout section is checked against. Run in the build container, where the
reference is readable; the fixture travels, the reference does not."""
import json
import os

SRC = "/root/reference/framework/tests/launch_config.json"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "launch_layout.json")

if __name__ == "__main__":
    with open(SRC, encoding="utf-8") as f:
        layout = json.load(f)["layout"]
    with open(OUT, "w", encoding="utf-8") as f:
        json.dump(layout, f, indent=1)
    print(f"wrote {OUT}: {layout}")
