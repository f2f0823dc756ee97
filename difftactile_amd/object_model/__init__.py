"""Drop-in MPM object models (``difftactile/object_model/``) on the HIP library."""
