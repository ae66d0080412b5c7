"""ctypes view of the reference's plugin interface
(include/basements/engine/plugins/plugin_capi.h:34-62): what an engine sees
after dlopen of a plugin library and a call of basements_plugin_register."""
from __future__ import annotations

import ctypes as C


class PluginInfo(C.Structure):
    _fields_ = [("name", C.c_char_p), ("version", C.c_char_p), ("author", C.c_char_p),
                ("description", C.c_char_p), ("api_version", C.c_uint32)]


class PluginCallbacks(C.Structure):
    _fields_ = [("get_info", C.c_void_p), ("on_load", C.c_void_p), ("on_unload", C.c_void_p),
                ("on_pre_step", C.c_void_p), ("on_post_step", C.c_void_p), ("on_render", C.c_void_p)]


class Plugin:
    """The plugin's callbacks and its own exports, as an engine would bind them."""

    def __init__(self, path):
        self.dll = C.CDLL(path)
        reg = self.dll.basements_plugin_register
        reg.restype = PluginCallbacks
        reg.argtypes = []
        self.cb = reg()
        self.get_info = C.CFUNCTYPE(PluginInfo)(self.cb.get_info)
        self.on_load = C.CFUNCTYPE(C.c_int, C.c_void_p)(self.cb.on_load)
        self.on_unload = C.CFUNCTYPE(None)(self.cb.on_unload)
        self.on_pre_step = C.CFUNCTYPE(None, C.c_float)(self.cb.on_pre_step)
        self.on_post_step = C.CFUNCTYPE(None, C.c_float)(self.cb.on_post_step)
        self.handle = self.dll.bm_soil_plugin_handle
        self.handle.restype = C.c_void_p
        self.handle.argtypes = []
        self.status = self.dll.bm_soil_plugin_status
        self.status.restype = C.c_int
        self.status.argtypes = []
        self.set_body = self.dll.bm_soil_plugin_set_body
        self.set_body.restype = C.c_int
        self.set_body.argtypes = [C.c_void_p]
