"""Learned starting density (filled in by the MLP guess module)."""


def load_default():
    return None
