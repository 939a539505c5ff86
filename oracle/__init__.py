"""CPU oracle of the parcel step -- test infrastructure only (see st_oracle.h)."""
