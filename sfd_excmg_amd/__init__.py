"""sfd_excmg_amd -- MI355X-native 3-D MT forward solve (SFD curl-curl + EXCMG)."""
