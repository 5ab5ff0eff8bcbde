"""Solver code strings, patch groups and their use by the grid reader
(codestring.cpp:33-212, patchgroups.cpp:27-112, patchgrid.cpp:541-549).

The normalisation rules are pinned by the reference's own example in
codestring.cpp:42-44 ("  aa, 034  00\\n 70;5 " -> "aa 34 0 70 5") and by cases
worked out by hand from the rules stated there; the grouping cases are worked
out by hand from patchgroups.cpp:35-101 (first group of the same type that
agrees takes the patch; a "0" word agrees with any word; a group's "0" that a
patch fills is replaced in the group's codes)."""
import pytest

from drnum_amd import host


@pytest.mark.parametrize("text,expected", [
    ("  aa, 034  00\n 70;5 ", "aa 34 0 70 5"),      # codestring.cpp:42-44
    ("fx fy fz far far far far far far 0", "fx fy fz far far far far far far 0"),
    ("fx,fy;fz\nfar", "fx fy fz far"),              # , ; and newline are blanks
    ("   a    b   ", "a b"),                          # runs of blanks, ends trimmed
    ("007 0 00 000", "7 0 0 0"),                      # a word of zeros keeps one
    ("a0 0b a00b", "a0 b a00b"),                      # only zeros that start a word go
    ("0fx", "fx"),
    ("a\tb", "a\tb"),                                 # a tab is not a delimiter here
    ("", ""),
])
def test_codestring_normalisation(text, expected):
    assert host.codestring(text) == expected


@pytest.mark.parametrize("a,b,expected", [
    ("fx fy fz far", "fx fy fz far", (True, False, False, "fx fy fz far")),
    # a "0" of `a` takes the word of `b`: overruled
    ("fx 0 fz far", "fx fy fz far", (False, True, False, "fx fy fz far")),
    # a "0" of `b` takes the word of `a`: underruled
    ("fx fy fz far", "fx fy 0 far", (False, False, True, "fx fy fz far")),
    ("fx 0 fz far", "fx fy 0 far", (False, True, True, "fx fy fz far")),
    # different words, neither "0": no merge, all flags false
    ("fx fy fz jet", "fx fy fz far", (False, False, False, "")),
    # a "0" before a conflict does not survive the conflict
    ("0 fy fz jet", "fx fy fz far", (False, False, False, "")),
    # different word counts
    ("fx fy fz", "fx fy fz far", (False, False, False, "")),
    ("fx fy fz 0", "fx fy fz", (False, False, False, "")),
    # both strings are normalised first
    ("fx, fy;fz", " fx fy  fz ", (True, False, False, "fx fy fz")),
])
def test_codestring_compare(a, b, expected):
    assert host.codestring_compare(a, b) == expected


def patch_entry(codes, x0=0.0):
    return ("1001 // comment\n{\n  %r 0 0\n  1 0 0\n  0 1 0\n  1\n  6 6 6\n  0 0 0 0 0 0\n  1 1 1\n  %s\n}\n"
            % (x0, codes))


def read_grid(tmp_path, codes_list):
    path = tmp_path / "grid"
    path.write_text("".join(patch_entry(c, 2.0 * n) for n, c in enumerate(codes_list)) + "0\n")
    g = host.HostGrid(True)
    g.read(str(path))
    return g


def test_patch_codes_are_read_and_normalised(tmp_path):
    g = read_grid(tmp_path, ["fx,fy;fz  far far\n far far far far 00", "fx fy fz"])
    assert g.patch_codes(0) == "fx fy fz far far far far far far 0"
    assert g.patch_codes(1) == "fx fy fz"


def test_patch_groups(tmp_path):
    codes = [
        "fx fy fz far 0 far far far far 0",     # 0: new group A
        "fx fy fz 0 far far far far far 0",     # 1: A, fills A's "0" -> A = "... far far far ..."
        "fx fy fz jet 0 far far far far 0",     # 2: "jet" against A's "far": new group B
        "fx fy fz far far far far far far 0",   # 3: equal to A's merged codes
        "fx fy fz far",                         # 4: other word count: new group C
        "fx fy fz 0 0 0 0 0 0 0",               # 5: agrees with A and B; the first group wins
    ]
    g = read_grid(tmp_path, codes)
    groups = g.patch_groups()
    assert [x["patches"] for x in groups] == [[0, 1, 3, 5], [2], [4]]
    assert [x["codes"] for x in groups] == ["fx fy fz far far far far far far 0",
                                            "fx fy fz jet 0 far far far far 0", "fx fy fz far"]
    assert [x["exceptions"] for x in groups] == [True, False, False]
    assert all(x["type"] == 1001 for x in groups)


def test_equal_codes_form_one_group_without_exceptions(tmp_path):
    g = read_grid(tmp_path, ["a b c 0"] * 3)
    assert g.patch_groups() == [dict(type=1001, patches=[0, 1, 2], codes="a b c 0", exceptions=False)]
    # iteratorfeeder.cpp:38-49: an iterator built for "a b c d" does not match such a group
    assert host.codestring("a b c d") != g.patch_groups()[0]["codes"]


def test_jet_grid_file_forms_one_group(tmp_path):
    """The jet grid file written as the reference's writer lays it out: faces
    with seek layers say "0", the others their functor; all 64 patches agree."""
    from casefiles import grid_file_text
    from drnum_amd import configs

    path = tmp_path / "grid"
    path.write_text(grid_file_text(configs.jet()))
    g = host.HostGrid(True)
    g.read(str(path))
    groups = g.patch_groups()
    assert len(groups) == 1
    assert groups[0]["patches"] == list(range(64))
    assert groups[0]["codes"] == "fx fy fz trans trans trans trans jet trans 0"
    assert groups[0]["exceptions"] is True
