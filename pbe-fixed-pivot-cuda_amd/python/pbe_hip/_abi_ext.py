"""ctypes signatures of the C-ABI entry points beyond the two reference
replacements (added as the library grows)."""


def declare(lib):
    return lib
