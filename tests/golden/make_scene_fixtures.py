"""Copy the reference's sample scenes (sample/*.xml, sample/cube.obj) into tests/golden/scenes
as test fixtures, with their comments removed (the loader tests need only the data).

    python tests/golden/make_scene_fixtures.py [/path/to/reference]
"""
import os
import re
import sys

SRC = os.path.join(sys.argv[1] if len(sys.argv) > 1 else "/root/reference", "sample")
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "scenes")


def main():
    os.makedirs(DST, exist_ok=True)
    for name in sorted(os.listdir(SRC)):
        path = os.path.join(SRC, name)
        text = open(path, encoding="utf-8").read()
        if name.endswith(".xml"):
            text = re.sub(r"<!--.*?-->", "", text, flags=re.S)
        elif name.endswith(".obj"):
            text = "".join(l for l in text.splitlines(keepends=True) if not l.lstrip().startswith("#"))
        else:
            continue
        text = "\n".join(l.rstrip() for l in text.splitlines() if l.strip()) + "\n"
        with open(os.path.join(DST, name), "w", encoding="utf-8") as f:
            f.write(text)
        print("wrote", name)


if __name__ == "__main__":
    main()
