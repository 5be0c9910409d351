"""Python side of the MI355X flowVC advection path.

`host`   - ctypes bindings of libfvc_host.so (decks, formats, synthetic inputs)
`deck`   - ompVC-compatible parameter decks
`fvc`    - ctypes bindings of libfvc.so (the C ABI of include/fvc.h; GPU)
"""
import os

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REPO_DIR = os.path.dirname(PKG_DIR)
