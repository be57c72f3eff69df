"""MI355X-native implementation of MaoMi's optical stage.

Importing the package loads nothing native; the HIP library
(maomi_amd/libmaomi_amd.so) is opened on the first GPU call.
"""

__version__ = "0.1.0"
