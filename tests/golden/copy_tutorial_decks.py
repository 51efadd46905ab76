"""Copies the reference's tutorial decks that the static hot-path subset reads
unchanged into this folder, as input fixtures (data: the decks carry no
expected outputs; tests/test_deck.py pins tutorial01's result by an
independent Reissner-beam integration and the unloaded end states of
tutorial07 and tutorial08).  Run where the reference checkout is
present:  python tests/golden/copy_tutorial_decks.py [reference root]"""
import os
import shutil
import sys

DECKS = [
    "tutorial01",  # Beam_1 cantilever, Rectangle section, dead tip load, one Static step
    "tutorial03",  # Pipe_1 column, perturbation then follower compression, two Static steps
    "tutorial07",  # Beam_1 cantilever, loads and gravity switched on and off over six Static steps
    "tutorial08",  # Beam_1 cantilever, three Static steps ending unloaded (ConcomitantSolution skipped)
    "tutorial09",  # Beam_1 arc, two Static steps
]


def main(root="/root/reference"):
    here = os.path.dirname(os.path.abspath(__file__))
    for name in DECKS:
        src = os.path.join(root, "inputs", name, name + ".inp")
        shutil.copyfile(src, os.path.join(here, name + ".inp"))
        print("copied", src)


if __name__ == "__main__":
    main(*sys.argv[1:])
